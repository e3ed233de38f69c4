// Batched CG-SENSE: many independent reconstructions of one image / grid
// shape (slices with uneven spoke and coil counts, or frames sharing one
// packed trajectory) advance in lockstep, each step one launch for all of
// them: the FFT passes put the problem in blockIdx.z, the interpolation
// merges every problem's work units into one launch (units of a shared pack
// run unit-major so its tiles stream from HBM once), and the CG vector
// kernels put the problem in blockIdx.y.  Scalars stay per problem (float64).
//
// Semantics per problem are exactly tsn_cg's (same kernels, same reduction
// order), so a batched solve reproduces the single-plan solve bit for bit.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "cg_kernels.h"

namespace tsn {
int launch_split_panel_il(const float2* in, uint32_t rows, uint32_t coils, float scale, uint32_t* out,
                          hipStream_t stream, int values);
}  // namespace tsn

using namespace tsn;

namespace {

// measurement z (coil-major [coils][m_z], raw) * scale -> interleaved split
// sample panel of problem z (blocks of [m_z][16]).
template <bool BF16>
__global__ __launch_bounds__(256) void batch_split_kernel(const float2* const* __restrict__ meas,
                                                          const uint32_t* __restrict__ coils_of,
                                                          const uint32_t* __restrict__ m_of, uint32_t blocks_max,
                                                          float scale, uint32_t* __restrict__ panels,
                                                          uint64_t panel_stride) {
  const uint32_t z = blockIdx.y;
  const uint32_t m = m_of[z], coils = coils_of[z];
  const uint32_t blocks = (coils + kCoilsPerBlock - 1) / kCoilsPerBlock;
  const uint64_t total = static_cast<uint64_t>(blocks) * m * kN;
  uint32_t* out = panels + z * panel_stride;
  for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const int n = static_cast<int>(g & 15);
    const uint64_t br = g >> 4;
    const uint32_t row = static_cast<uint32_t>(br % m), b = static_cast<uint32_t>(br / m);
    const uint32_t coil = b * kCoilsPerBlock + (n & 7);
    float v = 0.f;
    if (coil < coils) {
      const float2 x = meas[z][static_cast<size_t>(coil) * m + row];
      v = (n < 8 ? x.x : x.y) * scale;
    }
    uint32_t w;
    if constexpr (BF16) {
      w = split_bf16(v);
    } else {
      const _Float16 h = static_cast<_Float16>(v);
      const _Float16 r = static_cast<_Float16>(v - static_cast<float>(h));
      w = static_cast<uint32_t>(__builtin_bit_cast(uint16_t, h)) |
          (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, r)) << 16);
    }
    out[g] = w;
  }
  (void)blocks_max;
}

// x_z (column-major [n1][n0]) -> x_out[z] ([n0][n1]); history sqrt(rho) and
// status out.  A launch of its own after cg_final_kernel, so every block of
// that kernel has read rho before the history is read here.
__global__ __launch_bounds__(256) void batch_out_kernel(const float2* __restrict__ x, uint64_t vstride,
                                                        uint32_t n0, uint32_t n1, float2* const* __restrict__ xout,
                                                        const double* __restrict__ hist, uint32_t hstride,
                                                        uint32_t iterations, double* const* __restrict__ hist_out,
                                                        const int* __restrict__ status, int* __restrict__ status_out) {
  const uint32_t z = blockIdx.y;
  const uint64_t px = static_cast<uint64_t>(n0) * n1;
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < px;
       e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t i1 = static_cast<uint32_t>(e % n1), i0 = static_cast<uint32_t>(e / n1);
    xout[z][e] = x[z * vstride + static_cast<uint64_t>(i1) * n0 + i0];
  }
  if (blockIdx.x == 0) {
    for (uint32_t i = threadIdx.x; i <= iterations; i += blockDim.x) hist_out[z][i] = sqrt(fmax(hist[z * hstride + i], 0.0));
    if (threadIdx.x == 0) status_out[z] = status[z];
  }
}

}  // namespace

struct tsn_batch {
  tsn_plan_config cfg{};
  int device = 0;
  uint32_t problems = 0, cmax = 0, bmax = 0, mmax = 0, pixels = 0, vec_blocks = 0, grid = 0;
  std::vector<uint32_t> coils, m;
  std::vector<const tsn_direction*> fwd, adj;
  InterpBatch ib_fwd[2], ib_adj[2];  // parts of one and of two coil blocks
  std::vector<InterpBatchPart> parts_fwd[2], parts_adj[2];  // (kept for the autotuner's rebuilds)
  uint32_t member_group = 0, coil_waves = kDefaultCoilWaves, depth = 0;
  std::vector<void*> allocations;
  size_t bytes = 0;
  float2 *tw0 = nullptr, *tw1 = nullptr, *smaps_t = nullptr, *t_fwd = nullptr, *t_inv = nullptr, *grid_buf = nullptr;
  float* apod_t = nullptr;
  uint32_t *pan_il = nullptr, *smp_il = nullptr, *coils_dev = nullptr, *m_dev = nullptr;
  float2 *x = nullptr, *r = nullptr, *r2 = nullptr, *p = nullptr, *p2 = nullptr, *ap = nullptr;
  double *partial = nullptr, *hist = nullptr;
  uint32_t hist_cap = 0;
  int* status = nullptr;
  // per-call tables (device): measurement / output pointers and inverse-pass scales
  const float2** meas_tab = nullptr;
  float2** xout_tab = nullptr;
  double** hist_tab = nullptr;
  float2 *scale_rhs = nullptr, *scale_normal = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

template <typename T>
int batch_alloc(tsn_batch* bt, T** ptr, size_t count) {
  *ptr = nullptr;
  if (count == 0) return TSN_OK;
  TSN_HIP(hipMalloc(reinterpret_cast<void**>(ptr), count * sizeof(T)));
  bt->allocations.push_back(*ptr);
  bt->bytes += count * sizeof(T);
  return TSN_OK;
}

void batch_destroy(tsn_batch* bt) {
  if (!bt) return;
  if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
  if (bt->graph) (void)hipGraphDestroy(bt->graph);
  for (int k = 0; k < 2; ++k) {
    interp_batch_free(&bt->ib_fwd[k]);
    interp_batch_free(&bt->ib_adj[k]);
  }
  for (void* q : bt->allocations) (void)hipFree(q);
  delete bt;
}

int interp_kind_of(int precision) {
  switch (precision) {
    case TSN_PREC_DENSE16: return 1;
    case TSN_PREC_HI16: return 2;
    case TSN_PREC_SPLITBF16: return 3;
    default: return 0;
  }
}

uint64_t panel_stride(const tsn_batch* bt) { return static_cast<uint64_t>(bt->bmax) * bt->grid * kN; }
uint64_t sample_stride(const tsn_batch* bt) { return static_cast<uint64_t>(bt->bmax) * bt->mmax * kN; }

FftArgs batch_fft_args(const tsn_batch* bt) {
  FftArgs a{};
  a.n0 = bt->cfg.image0;
  a.n1 = bt->cfg.image1;
  a.k0 = bt->cfg.grid0;
  a.k1 = bt->cfg.grid1;
  a.coils = bt->cmax;
  a.cstride = bt->cmax;
  a.coils_of = bt->coils_dev;
  a.tw0 = bt->tw0;
  a.tw1 = bt->tw1;
  a.smaps_t = bt->smaps_t;
  a.apod_t = bt->apod_t;
  a.t_fwd = bt->t_fwd;
  a.panel_il = bt->pan_il;
  a.grid = bt->grid_buf;
  a.t_inv = bt->t_inv;
  a.grid_blocked = 1;  // each problem's rows compact (<= 8 coils) or in 8-coil blocks
  a.image_stride = bt->pixels;
  a.smaps_stride = static_cast<uint64_t>(bt->pixels) * bt->cmax;
  a.t_stride = static_cast<uint64_t>(bt->cfg.grid0) * bt->cfg.image1 * bt->bmax * kCoilsPerBlock;
  a.grid_stride = static_cast<uint64_t>(bt->grid) * bt->bmax * kCoilsPerBlock;
  a.panel_stride = panel_stride(bt);
  a.partial_stride = 3 * static_cast<uint64_t>(bt->cfg.image1);
  a.hist_stride = bt->hist_cap;
  return a;
}

int out_kind_of(int precision) { return precision == TSN_PREC_SPLITBF16 ? 2 : 0; }

int enqueue_batch_cg(tsn_batch* bt, uint32_t iterations, hipStream_t s) {
  const uint32_t S = bt->problems;
  const int kind = interp_kind_of(bt->cfg.precision);
  const float norm = static_cast<float>(1.0 / std::sqrt(static_cast<double>(bt->cfg.grid0) * bt->cfg.grid1));
  const uint32_t n = bt->pixels, vb = bt->vec_blocks;
  const uint64_t vs = bt->pixels;
  // b = A^H (y / sqrt(K0K1)) / s^2 per problem
  const uint64_t total = static_cast<uint64_t>(bt->bmax) * bt->mmax * kN;
  const dim3 split_grid(std::min<uint64_t>(div_up(total, 256), 4096), S);
  if (bt->cfg.precision == TSN_PREC_SPLITBF16)
    hipLaunchKernelGGL(batch_split_kernel<true>, split_grid, dim3(256), 0, s, bt->meas_tab, bt->coils_dev,
                       bt->m_dev, bt->bmax, norm, bt->smp_il, sample_stride(bt));
  else
    hipLaunchKernelGGL(batch_split_kernel<false>, split_grid, dim3(256), 0, s, bt->meas_tab, bt->coils_dev,
                       bt->m_dev, bt->bmax, norm, bt->smp_il, sample_stride(bt));
  for (const InterpBatch& ib : bt->ib_adj)
    if (ib.n_problems) TSN_CHECK(launch_interp_batch(ib, kind, 1, true, 1.0f, s));
  // b written straight into p (r0 = p0 = b, x0 = 0), with the per-column
  // partials of <b, b> in the <v, v> slot; the plan's schedule (plan.hip
  // enqueue_cg) for every problem at once
  const uint32_t cols = bt->cfg.image1;
  {
    FftArgs a = batch_fft_args(bt);
    a.image_out = bt->p;
    a.scale_of = bt->scale_rhs;
    a.dot_partial = bt->partial;
    a.dot_r = bt->p;
    TSN_CHECK(fft_inverse(a, s, S));
  }
  float2* p_cur = bt->p;
  float2* p_other = bt->p2;
  float2* r_cur = bt->p;
  float2* r_other = bt->r2;
  for (uint32_t it = 0; it < iterations; ++it) {
    FftArgs a = batch_fft_args(bt);
    a.rho_hist = bt->hist;
    a.cg_status = bt->status;
    a.dot_count = cols;
    if (it == 0) {
      a.image_t = p_cur;
      a.rho0_dots = bt->partial + 2 * cols;
    } else {
      a.cg_r = r_cur;
      a.cg_r_new = r_other;
      a.cg_x = bt->x;
      a.x_zero = it == 1;
      a.cg_ap = bt->ap;
      a.cg_p_old = p_cur;
      a.cg_p_new = p_other;
      a.cg_dots = bt->partial;
      a.iter = it;
      std::swap(p_cur, p_other);
      float2* const r_done = r_cur;
      r_cur = r_other;
      r_other = it == 1 ? bt->r : r_done;  // r_0 was p_0's buffer, now p_other
    }
    TSN_CHECK(fft_forward(a, out_kind_of(bt->cfg.precision), s, S));
    for (const InterpBatch& ib : bt->ib_fwd)
      if (ib.n_problems) TSN_CHECK(launch_interp_batch(ib, kind, 3, false, norm, s));
    for (const InterpBatch& ib : bt->ib_adj)
      if (ib.n_problems) TSN_CHECK(launch_interp_batch(ib, kind, 1, true, 1.0f, s));
    FftArgs inv = batch_fft_args(bt);
    inv.image_out = bt->ap;
    inv.reg_in = p_cur;
    inv.scale_of = bt->scale_normal;
    inv.dot_partial = bt->partial;
    inv.dot_r = r_cur;
    TSN_CHECK(fft_inverse(inv, s, S));
  }
  hipLaunchKernelGGL(cg_final_kernel, dim3(vb, S), dim3(kVecThreads), 0, s, bt->x, p_cur, bt->partial,
                     bt->cfg.image1, bt->hist, iterations - 1, iterations == 1 ? 1 : 0, bt->status, n, vs,
                     3 * bt->cfg.image1, bt->hist_cap);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

int enqueue_batch_out(tsn_batch* bt, uint32_t iterations, int* status_out, hipStream_t s) {
  hipLaunchKernelGGL(batch_out_kernel, dim3(std::min<uint32_t>(div_up(bt->pixels, 256), 1024), bt->problems),
                     dim3(256), 0, s, bt->x, static_cast<uint64_t>(bt->pixels), bt->cfg.image0, bt->cfg.image1,
                     bt->xout_tab, bt->hist, bt->hist_cap, iterations, bt->hist_tab, bt->status, status_out);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

// Upload the per-call tables (synchronously: the caller's stream is drained
// first, so no queued launch still reads the previous tables).
int stage_call(tsn_batch* bt, const void* const* meas, uint32_t iterations, const float* scales, float lambda,
               void* const* x_out, void* const* residuals, hipStream_t s) {
  const uint32_t S = bt->problems;
  TSN_HIP(hipSetDevice(bt->device));
  TSN_HIP(hipStreamSynchronize(s));
  if (iterations + 1 > bt->hist_cap) {
    if (bt->graph) return fail(TSN_ERR_ARGUMENT, "iterations exceed the captured batch's history capacity");
    double* h = nullptr;
    TSN_HIP(hipMalloc(reinterpret_cast<void**>(&h), static_cast<size_t>(S) * (iterations + 1) * sizeof(double)));
    bt->allocations.push_back(h);
    bt->hist = h;
    bt->hist_cap = iterations + 1;
  }
  const double kk = static_cast<double>(bt->cfg.grid0) * bt->cfg.grid1;
  const float norm = static_cast<float>(1.0 / std::sqrt(kk));
  std::vector<float2> rhs(S), nrm(S);
  for (uint32_t z = 0; z < S; ++z) {
    const double sq = static_cast<double>(scales[z]) * scales[z];
    const float inv = static_cast<float>(1.0 / sq);
    rhs[z] = make_float2(inv * norm, 0.0f);
    nrm[z] = make_float2(inv * norm, static_cast<float>(lambda / kk / sq));
  }
  TSN_HIP(hipMemcpy(bt->meas_tab, meas, S * sizeof(void*), hipMemcpyHostToDevice));
  TSN_HIP(hipMemcpy(bt->xout_tab, x_out, S * sizeof(void*), hipMemcpyHostToDevice));
  TSN_HIP(hipMemcpy(bt->hist_tab, residuals, S * sizeof(void*), hipMemcpyHostToDevice));
  TSN_HIP(hipMemcpy(bt->scale_rhs, rhs.data(), S * sizeof(float2), hipMemcpyHostToDevice));
  TSN_HIP(hipMemcpy(bt->scale_normal, nrm.data(), S * sizeof(float2), hipMemcpyHostToDevice));
  return TSN_OK;
}

int check_batch_call(tsn_batch* bt, const void* const* meas, uint32_t iterations, const float* scales,
                     float lambda, void* const* x_out, void* const* residuals, const void* status) {
  if (!bt || !meas || !scales || !x_out || !residuals || !status) return fail(TSN_ERR_ARGUMENT, "null argument");
  if (iterations == 0) return fail(TSN_ERR_ARGUMENT, "cg iterations must be positive");
  if (!(lambda >= 0.0f) || !std::isfinite(lambda)) return fail(TSN_ERR_ARGUMENT, "lambda must be >= 0");
  for (uint32_t z = 0; z < bt->problems; ++z) {
    if (!meas[z] || !x_out[z] || !residuals[z]) return fail(TSN_ERR_ARGUMENT, "null pointer for problem %u", z);
    if (!(scales[z] > 0.0f) || !std::isfinite(scales[z]))
      return fail(TSN_ERR_ARGUMENT, "operator scale of problem %u must be positive and finite", z);
  }
  return TSN_OK;
}

// (Re)build the batched interpolation tables with the batch's member group
// and coil waves.
int build_interp_tables(tsn_batch* bt) {
  for (int k = 0; k < 2; ++k) {
    interp_batch_free(&bt->ib_fwd[k]);
    interp_batch_free(&bt->ib_adj[k]);
  }
  for (int k = 0; k < 2; ++k) {
    if (bt->parts_fwd[k].empty()) continue;
    TSN_CHECK(interp_batch_build(bt->parts_fwd[k], &bt->ib_fwd[k], bt->member_group, bt->coil_waves, bt->depth));
    TSN_CHECK(interp_batch_build(bt->parts_adj[k], &bt->ib_adj[k], bt->member_group, bt->coil_waves, bt->depth));
  }
  return TSN_OK;
}

// One forward and one adjoint interpolation step of every problem.
int enqueue_interp_steps(tsn_batch* bt, hipStream_t s) {
  const int kind = interp_kind_of(bt->cfg.precision);
  const float norm = static_cast<float>(1.0 / std::sqrt(static_cast<double>(bt->cfg.grid0) * bt->cfg.grid1));
  for (const InterpBatch& ib : bt->ib_fwd)
    if (ib.n_problems) TSN_CHECK(launch_interp_batch(ib, kind, 3, false, norm, s));
  for (const InterpBatch& ib : bt->ib_adj)
    if (ib.n_problems) TSN_CHECK(launch_interp_batch(ib, kind, 1, true, 1.0f, s));
  return TSN_OK;
}

}  // namespace

extern "C" {

int tsn_batch_autotune(tsn_batch* bt, void* stream, uint32_t* best_member_group, uint32_t* best_coil_waves,
                       uint32_t* best_depth, float* best_us) {
  if (!bt) return fail(TSN_ERR_ARGUMENT, "null batch");
  if (bt->graph) return fail(TSN_ERR_ARGUMENT, "tune a batch before capturing its graph");
  hipStream_t s = static_cast<hipStream_t>(stream);
  TSN_HIP(hipSetDevice(bt->device));
  TSN_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> groups;
  for (uint32_t g : {1u, 2u, 4u, 8u, 16u, bt->problems})
    if (g <= bt->problems && std::find(groups.begin(), groups.end(), g) == groups.end()) groups.push_back(g);
  std::vector<uint32_t> waves = {bt->coil_waves};
  if (!bt->parts_fwd[1].empty()) waves = {1u, 2u};
  // pipeline depths: the kernels' default and every alternative (one-block
  // launches take 1 or 3, two-block launches 2; others keep their default)
  const std::vector<uint32_t> depths = {0u, 1u, 2u, 3u};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = TSN_OK;
  auto hip = [&](hipError_t e, const char* what) {
    if (rc == TSN_OK && e != hipSuccess) rc = fail(TSN_ERR_RUNTIME, "batch autotune %s: %s", what, hipGetErrorString(e));
  };
  hip(hipEventCreate(&e0), "event");
  hip(hipEventCreate(&e1), "event");
  // two passes over the candidates, each candidate's faster pass counts
  const size_t n_cand = waves.size() * groups.size() * depths.size();
  std::vector<float> times(n_cand, 0.0f);
  uint32_t best_g = 0, best_w = 0, best_d = 0;
  float best_ms = 0.0f;
  for (int pass = 0; pass < 2; ++pass)
  for (size_t ci = 0; ci < n_cand; ++ci) {
    if (rc != TSN_OK) break;
    const uint32_t w = waves[ci / (groups.size() * depths.size())];
    const uint32_t g = groups[(ci / depths.size()) % groups.size()];
    const uint32_t dp = depths[ci % depths.size()];
    bt->member_group = g;
    bt->coil_waves = w;
    bt->depth = dp;
    rc = build_interp_tables(bt);
    for (int i = 0; rc == TSN_OK && i < 2; ++i) rc = enqueue_interp_steps(bt, s);
    if (rc != TSN_OK) break;
    hip(hipEventRecord(e0, s), "event");
    constexpr int kReps = 4;
    for (int i = 0; rc == TSN_OK && i < kReps; ++i) rc = enqueue_interp_steps(bt, s);
    hip(hipEventRecord(e1, s), "event");
    hip(hipEventSynchronize(e1), "event");
    float ms = 0.0f;
    hip(hipEventElapsedTime(&ms, e0, e1), "event");
    ms /= kReps;
    if (rc != TSN_OK) break;
    if (pass == 1) ms = std::min(ms, times[ci]);
    times[ci] = ms;
    if (pass == 1 && (best_g == 0 || ms < best_ms)) {
      best_g = g;
      best_w = w;
      best_d = dp;
      best_ms = ms;
    }
  }
  (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc != TSN_OK) return rc;
  bt->member_group = best_g;
  bt->coil_waves = best_w;
  bt->depth = best_d;
  TSN_CHECK(build_interp_tables(bt));
  TSN_HIP(hipStreamSynchronize(s));
  if (best_member_group) *best_member_group = best_g;
  if (best_coil_waves) *best_coil_waves = best_w;
  if (best_depth) *best_depth = best_d;
  if (best_us) *best_us = best_ms * 1000.0f;
  return TSN_OK;
}

int tsn_batch_create(const tsn_plan_config* cfg, const tsn_batch_problem* problems, uint32_t count,
                     const void* apodization, tsn_batch** out) {
  if (!cfg || !problems || !apodization || !out || count == 0) return fail(TSN_ERR_ARGUMENT, "bad arguments");
  *out = nullptr;
  const tsn_plan_config c = *cfg;
  if (c.grid0 != 2 * c.image0 || c.grid1 != 2 * c.image1 || !fft_length_supported(c.grid0) ||
      !fft_length_supported(c.grid1))
    return fail(TSN_ERR_ARGUMENT, "grid must be 2x image with grid sides in {64, 128, 256, 512, 768}");
  if (c.precision == TSN_PREC_FP32 || c.precision < TSN_PREC_SPLIT16 || c.precision > TSN_PREC_SPLITBF16)
    return fail(TSN_ERR_ARGUMENT, "batches run the packed variants (split16, dense16, hi16, splitbf16)");
  const int want = c.precision == TSN_PREC_SPLITBF16 ? TSN_VALUES_BF16 : TSN_VALUES_FP16;
  const uint32_t grid = c.grid0 * c.grid1;
  auto* bt = new tsn_batch();
  bt->cfg = c;
  (void)hipGetDevice(&bt->device);
  bt->problems = count;
  bt->grid = grid;
  bt->pixels = c.image0 * c.image1;
  bt->vec_blocks = div_up(bt->pixels, kVecThreads);
  for (uint32_t z = 0; z < count; ++z) {
    const tsn_batch_problem& q = problems[z];
    int rc = TSN_OK;
    if (!q.forward || !q.adjoint || !q.smaps) rc = fail(TSN_ERR_ARGUMENT, "problem %u: null argument", z);
    else if (q.coils == 0 || q.coils > static_cast<uint32_t>(kMaxCoils))
      rc = fail(TSN_ERR_ARGUMENT, "problem %u: coils must be in 1..16", z);
    else if (q.forward->csr || q.adjoint->csr || q.forward->values != want || q.adjoint->values != want)
      rc = fail(TSN_ERR_ARGUMENT, "problem %u: packs do not match precision %d", z, c.precision);
    else if (q.forward->h.n_cols != grid || q.adjoint->h.output_size != grid ||
             q.forward->h.output_size != q.adjoint->h.n_cols)
      rc = fail(TSN_ERR_INCOMPATIBLE, "problem %u: unexpected operator shape", z);
    else if (c.precision == TSN_PREC_DENSE16 &&
             (ensure_dense_tiles(q.forward) != TSN_OK || ensure_dense_tiles(q.adjoint) != TSN_OK)) {
      const std::string why = tsn_last_error();
      rc = fail(TSN_ERR_ARGUMENT, "problem %u: dense-control tiles: %s", z, why.c_str());
    }
    if (rc != TSN_OK) {
      batch_destroy(bt);
      return rc;
    }
    bt->coils.push_back(q.coils);
    bt->m.push_back(q.forward->h.output_size);
    bt->fwd.push_back(q.forward);
    bt->adj.push_back(q.adjoint);
    bt->cmax = std::max(bt->cmax, q.coils);
    bt->mmax = std::max(bt->mmax, q.forward->h.output_size);
  }
  bt->bmax = div_up(bt->cmax, kCoilsPerBlock);
  const size_t px = bt->pixels, S = count;
  int rc = TSN_OK;
  auto A = [&](auto** ptr, size_t n) {
    if (rc == TSN_OK) rc = batch_alloc(bt, ptr, n);
  };
  A(&bt->tw0, c.grid0);
  A(&bt->tw1, c.grid1);
  A(&bt->smaps_t, S * px * bt->cmax);
  A(&bt->apod_t, px);
  // T arrays and grids: per problem, compact rows of its coils (<= 8) or 64-byte
  // rows in 8-coil blocks, at a common problem stride
  const size_t slots = static_cast<size_t>(bt->bmax) * kCoilsPerBlock;
  A(&bt->t_fwd, S * c.grid0 * c.image1 * slots);
  A(&bt->t_inv, S * c.grid0 * c.image1 * slots);
  A(&bt->grid_buf, S * grid * slots);
  A(&bt->pan_il, S * panel_stride(bt));
  A(&bt->smp_il, S * sample_stride(bt));
  for (float2** v : {&bt->x, &bt->r, &bt->r2, &bt->p, &bt->p2, &bt->ap}) A(v, S * px);
  A(&bt->partial, S * 3 * static_cast<size_t>(c.image1));
  A(&bt->status, S);
  A(&bt->coils_dev, S);
  A(&bt->m_dev, S);
  A(&bt->meas_tab, S);
  A(&bt->xout_tab, S);
  A(&bt->hist_tab, S);
  A(&bt->scale_rhs, S);
  A(&bt->scale_normal, S);
  if (rc == TSN_OK) {
    const auto t0 = twiddles(c.grid0), t1 = twiddles(c.grid1);
    if (hipMemcpy(bt->tw0, t0.data(), t0.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(bt->tw1, t1.data(), t1.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(bt->coils_dev, bt->coils.data(), S * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(bt->m_dev, bt->m.data(), S * 4, hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(TSN_ERR_RUNTIME, "batch setup upload failed");
  }
  if (rc == TSN_OK) {
    rc = zero_async(bt->grid_buf, S * grid * slots * sizeof(float2), 0);
    if (rc == TSN_OK) rc = zero_async(bt->smaps_t, S * px * bt->cmax * sizeof(float2), 0);
  }
  for (uint32_t z = 0; rc == TSN_OK && z < count; ++z) {
    const uint64_t ns = static_cast<uint64_t>(bt->coils[z]) * px;
    hipLaunchKernelGGL(smaps_transpose_kernel, dim3(div_up(ns, 256)), dim3(256), 0, 0,
                       static_cast<const float2*>(problems[z].smaps), bt->coils[z], c.image0, c.image1, bt->cmax,
                       bt->smaps_t + z * px * bt->cmax);
  }
  if (rc == TSN_OK) {
    hipLaunchKernelGGL(apod_transpose_kernel, dim3(div_up(px, 256)), dim3(256), 0, 0,
                       static_cast<const float*>(apodization), c.image0, c.image1, bt->apod_t);
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = fail(TSN_ERR_RUNTIME, "batch setup: %s", hipGetErrorString(e));
  }
  // interpolation parts: one per problem, in the batch of its coil-block
  // count (block 1's panels follow block 0's in pan_il and smp_il)
  if (rc == TSN_OK) {
    std::vector<InterpBatchPart> fw[2], ad[2];
    for (uint32_t z = 0; z < count; ++z) {
      const uint32_t blocks = div_up(bt->coils[z], kCoilsPerBlock);
      uint32_t* pan = bt->pan_il + z * panel_stride(bt);
      uint32_t* smp = bt->smp_il + z * sample_stride(bt);
      float* grid_out = reinterpret_cast<float*>(bt->grid_buf + z * grid * slots);
      fw[blocks - 1].push_back({bt->fwd[z], pan, nullptr, smp, bt->coils[z], 0, blocks});
      ad[blocks - 1].push_back({bt->adj[z], smp, grid_out, nullptr, bt->coils[z], 0, blocks});
    }
    for (int k = 0; k < 2; ++k) {
      bt->parts_fwd[k] = fw[k];
      bt->parts_adj[k] = ad[k];
    }
    rc = build_interp_tables(bt);
  }
  if (rc != TSN_OK) {
    batch_destroy(bt);
    return rc;
  }
  *out = bt;
  return TSN_OK;
}

void tsn_batch_free(tsn_batch* bt) { batch_destroy(bt); }

size_t tsn_batch_bytes(const tsn_batch* bt) { return bt ? bt->bytes : 0; }

int tsn_batch_cg(tsn_batch* bt, const void* const* measurements, uint32_t iterations, const float* operator_scales,
                 float lambda, void* const* x_out, void* const* residuals_out, void* status_out, void* stream) {
  TSN_CHECK(check_batch_call(bt, measurements, iterations, operator_scales, lambda, x_out, residuals_out,
                             status_out));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (bt->graph) return fail(TSN_ERR_ARGUMENT, "this batch holds a captured graph; launch it or free the batch");
  TSN_CHECK(stage_call(bt, measurements, iterations, operator_scales, lambda, x_out, residuals_out, s));
  TSN_CHECK(enqueue_batch_cg(bt, iterations, s));
  return enqueue_batch_out(bt, iterations, static_cast<int*>(status_out), s);
}

int tsn_batch_cg_graph_capture(tsn_batch* bt, const void* const* measurements, uint32_t iterations,
                               const float* operator_scales, float lambda, void* const* x_out,
                               void* const* residuals_out, void* status_out, void* stream) {
  TSN_CHECK(check_batch_call(bt, measurements, iterations, operator_scales, lambda, x_out, residuals_out,
                             status_out));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!s) return fail(TSN_ERR_ARGUMENT, "graph capture needs a non-null stream");
  if (bt->exec) {
    (void)hipGraphExecDestroy(bt->exec);
    bt->exec = nullptr;
  }
  if (bt->graph) {
    (void)hipGraphDestroy(bt->graph);
    bt->graph = nullptr;
  }
  TSN_CHECK(stage_call(bt, measurements, iterations, operator_scales, lambda, x_out, residuals_out, s));
  TSN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
  int rc = enqueue_batch_cg(bt, iterations, s);
  if (rc == TSN_OK) rc = enqueue_batch_out(bt, iterations, static_cast<int*>(status_out), s);
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(s, &g);
  if (rc != TSN_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (end != hipSuccess) return fail(TSN_ERR_RUNTIME, "hipStreamEndCapture: %s", hipGetErrorString(end));
  bt->graph = g;
  TSN_HIP(hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0));
  return TSN_OK;
}

int tsn_batch_cg_graph_launch(tsn_batch* bt, void* stream) {
  if (!bt || !bt->exec) return fail(TSN_ERR_ARGUMENT, "no captured graph");
  TSN_HIP(hipGraphLaunch(bt->exec, static_cast<hipStream_t>(stream)));
  return TSN_OK;
}

}  // extern "C"
