// Reconstruction plan: operators A, A^H, A^H A and CG-SENSE with float64
// scalars, eager or captured as one hipGraph.
//
// Schedule and conventions follow the reference production binary
// (experiments/production/src/nufft_fp16x2_cg_bench.cu):
//   measurement / sqrt(K0*K1)                     :329-334
//   forward interpolation * 1/sqrt(K0*K1)         :230-236
//   inverse FFT * apod * sqrt(K0K1)/(K0K1)        :335-352 (prenormalised)
//   normal = combine * 1/s^2 + lambda/(K0K1)/s^2  :27-54, 401-404
//   CG update order, curvature flag (alpha = 0)   :56-180, 241-268
// Differences by design: float64 CG scalars (rho, <p,Ap>, alpha, beta, history),
// reduced in fixed order by every block that needs them, so an iteration is
// six kernels (the x, r and direction updates ride in the first forward FFT
// pass, with ||r - alpha Ap||^2 expanded from three inverse-pass partials);
// image vectors kept column-major [n1][n0] inside the plan; SENSE combine,
// Tikhonov term and the <p, Ap> partial fused into the last inverse FFT pass;
// the forward interpolation writes the adjoint's FP16 hi/res sample panels
// directly (no float2 sample round trip).
#include <cmath>
#include <mutex>

#include "fft_common.h"

namespace tsn {
int launch_split_panel_il(const float2* in, uint32_t rows, uint32_t coils, float scale, uint32_t* out,
                          hipStream_t stream, int values);
int launch_split_panel(const float2* in, uint32_t rows, uint32_t coils, float scale, _Float16* hi,
                       _Float16* res, hipStream_t stream);
}  // namespace tsn

using namespace tsn;

#include "cg_kernels.h"

struct tsn_plan {
  tsn_plan_config cfg{};
  int device = 0;
  const tsn_direction* fwd = nullptr;
  const tsn_direction* adj = nullptr;
  uint32_t m = 0, blocks = 0, pixels = 0, vec_blocks = 0;
  size_t bytes = 0;
  std::vector<void*> allocations;
  float2 *tw0 = nullptr, *tw1 = nullptr, *smaps_t = nullptr, *t_fwd = nullptr, *t_inv = nullptr;
  float* apod_t = nullptr;
  uint32_t *pan_il = nullptr, *smp_il = nullptr;  // interleaved hi | res << 16 panels
  float2* y_rows = nullptr;                         // measurement rows [M][coils] (data consistency)
  float2 *pan32 = nullptr, *smp32 = nullptr, *grid = nullptr;
  float2 *x = nullptr, *r = nullptr, *r2 = nullptr, *p = nullptr, *p2 = nullptr, *ap = nullptr, *tmp = nullptr;
  // [3][image1] partials of <p, Ap>, Re<r, Ap>, <Ap, Ap> per image column (first: <b, b> in the third)
  double* partial = nullptr;
  // chunk slots and counters of the two directions' long groups, the plan's
  // own (so plans sharing a direction can run on different streams)
  ChunkScratch fwd_chunks, adj_chunks;
  // rho_k of the running solve (squared residual norms), f64[rho_cap]; the
  // caller's history buffer only receives sqrt(rho_k) from the final kernel
  double* rho = nullptr;
  uint32_t rho_cap = 0;
  int* status = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// rho entries allocated with a plan (64 CG iterations); longer solves grow it
constexpr uint32_t kRhoInitial = 65;
constexpr uint32_t kMaxCgIterations = 1u << 20;

template <typename T>
int plan_alloc(tsn_plan* pl, T** ptr, size_t count) {
  *ptr = nullptr;
  if (count == 0) return TSN_OK;
  TSN_HIP(hipMalloc(reinterpret_cast<void**>(ptr), count * sizeof(T)));
  pl->allocations.push_back(*ptr);
  pl->bytes += count * sizeof(T);
  return TSN_OK;
}


FftArgs fft_args(const tsn_plan* pl) {
  FftArgs a{};
  a.n0 = pl->cfg.image0;
  a.n1 = pl->cfg.image1;
  a.k0 = pl->cfg.grid0;
  a.k1 = pl->cfg.grid1;
  a.coils = pl->cfg.coils;
  a.cstride = pl->cfg.coils;
  a.tw0 = pl->tw0;
  a.tw1 = pl->tw1;
  a.smaps_t = pl->smaps_t;
  a.apod_t = pl->apod_t;
  a.t_fwd = pl->t_fwd;
  a.panel_il = pl->pan_il;
  a.panel32 = pl->pan32;
  a.grid = pl->grid;
  a.t_inv = pl->t_inv;
  // the packed variants' adjoint writes the grid in the FFT's chunk layout;
  // the FP32 CSR kernels gather and write [row][coils] rows
  a.grid_blocked = pl->cfg.precision == TSN_PREC_FP32 ? 0u : 1u;
  return a;
}

// float2 per row of the T arrays (and of the packed variants' grid): a
// problem's own coils up to 8, else 8 per coil block
uint32_t chunk_slots(uint32_t coils) {
  constexpr uint32_t kB = kCoilsPerBlock;
  return coils <= kB ? coils : div_up(coils, kB) * kB;
}

bool is_fp32(const tsn_plan* pl) { return pl->cfg.precision == TSN_PREC_FP32; }
bool is_bf16(const tsn_plan* pl) { return pl->cfg.precision == TSN_PREC_SPLITBF16; }
// forward FFT output: 0 FP16 interleaved split panel, 1 FP32 rows, 2 BF16 interleaved split panel
int fft_out_kind(const tsn_plan* pl) { return is_fp32(pl) ? 1 : is_bf16(pl) ? 2 : 0; }

int interp_kind(const tsn_plan* pl) {
  switch (pl->cfg.precision) {
    case TSN_PREC_DENSE16: return 1;
    case TSN_PREC_HI16: return 2;
    case TSN_PREC_SPLITBF16: return 3;
    default: return 0;
  }
}

float grid_norm(const tsn_plan* pl) {
  return static_cast<float>(1.0 / std::sqrt(static_cast<double>(pl->cfg.grid0) * pl->cfg.grid1));
}

// p (column-major image) -> gridded panels
int op_forward_grid(tsn_plan* pl, const float2* image_t, hipStream_t s) {
  FftArgs a = fft_args(pl);
  a.image_t = image_t;
  return fft_forward(a, fft_out_kind(pl), s);
}

// panels -> samples: split panels (internal) or float2 output (coil-major)
int op_forward_samples(tsn_plan* pl, float2* out_coil_major, hipStream_t s) {
  const float scale = grid_norm(pl);
  if (is_fp32(pl)) {
    if (out_coil_major)
      return launch_interp_csr(pl->fwd, pl->pan32, pl->cfg.coils, 0, false, false, scale, out_coil_major, s,
                               &pl->fwd_chunks);
    return launch_interp_csr(pl->fwd, pl->pan32, pl->cfg.coils, 1, false, false, scale, pl->smp32, s,
                             &pl->fwd_chunks);
  }
  if (out_coil_major)
    return launch_interp(pl->fwd, interp_kind(pl), pl->pan_il, nullptr, pl->cfg.coils, 0, false, false, scale,
                         reinterpret_cast<float*>(out_coil_major), nullptr, nullptr, s, true, nullptr,
                         &pl->fwd_chunks);
  return launch_interp(pl->fwd, interp_kind(pl), pl->pan_il, nullptr, pl->cfg.coils, 3, false, false, scale,
                       nullptr, nullptr, nullptr, s, true, pl->smp_il, &pl->fwd_chunks);
}

// sample panels -> grid (conjugate transpose interpolation)
int op_adjoint_grid(tsn_plan* pl, hipStream_t s) {
  if (is_fp32(pl))
    return launch_interp_csr(pl->adj, pl->smp32, pl->cfg.coils, 1, true, false, 1.0f, pl->grid, s,
                             &pl->adj_chunks, true);
  // row-major (mode 1) up to 8 coils, 8-coil blocks of rows (mode 5) above
  return launch_interp(pl->adj, interp_kind(pl), pl->smp_il, nullptr, pl->cfg.coils,
                       pl->cfg.coils > static_cast<uint32_t>(kCoilsPerBlock) ? 5 : 1, true, false, 1.0f, reinterpret_cast<float*>(pl->grid),
                       nullptr, nullptr, s, true, nullptr, &pl->adj_chunks);
}

// grid -> image (column-major) with combine, scale, regulariser, dot partial
int op_adjoint_image(tsn_plan* pl, float2* out_t, const float2* reg, float lambda, float out_scale,
                     double* dot_partial, hipStream_t s, const float2* dot_r = nullptr) {
  FftArgs a = fft_args(pl);
  a.image_out = out_t;
  a.reg_in = reg;
  a.lambda = lambda;
  a.out_scale = out_scale * grid_norm(pl);
  a.dot_partial = dot_partial;
  a.dot_r = dot_r;
  return fft_inverse(a, s);
}

int op_normal(tsn_plan* pl, const float2* in_t, float2* out_t, float inv_scale_sq, float lambda,
              double* dot_partial, hipStream_t s) {
  TSN_CHECK(op_forward_grid(pl, in_t, s));
  TSN_CHECK(op_forward_samples(pl, nullptr, s));
  TSN_CHECK(op_adjoint_grid(pl, s));
  return op_adjoint_image(pl, out_t, in_t, lambda, inv_scale_sq, dot_partial, s);
}

// measurement (raw, coil-major) -> scaled sample panels of the adjoint input
int op_load_samples(tsn_plan* pl, const float2* samples, float scale, hipStream_t s) {
  if (is_fp32(pl)) {
    const uint64_t n = static_cast<uint64_t>(pl->cfg.coils) * pl->m;
    hipLaunchKernelGGL(samples_to_rows_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, samples,
                       pl->cfg.coils, pl->m, scale, pl->smp32);
    TSN_HIP(hipGetLastError());
    return TSN_OK;
  }
  return launch_split_panel_il(samples, pl->m, pl->cfg.coils, scale, pl->smp_il, s,
                               is_bf16(pl) ? TSN_VALUES_BF16 : TSN_VALUES_FP16);
}

int transpose_image(const float2* in, uint32_t rows, uint32_t cols, float scale, float2* out, hipStream_t s) {
  hipLaunchKernelGGL(transpose_kernel, dim3(div_up(cols, 32), div_up(rows, 32)), dim3(32, 8), 0, s, in,
                     rows, cols, scale, out);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

int enqueue_cg(tsn_plan* pl, const float2* measurement, uint32_t iterations, float operator_scale,
               float lambda, float2* x_out, double* history, int* status_out, hipStream_t s) {
  const double sq = static_cast<double>(operator_scale) * operator_scale;
  const float inv_scale_sq = static_cast<float>(1.0 / sq);
  const double kk = static_cast<double>(pl->cfg.grid0) * pl->cfg.grid1;
  const float scaled_lambda = static_cast<float>(lambda / kk / sq);
  // b = A^H (y / sqrt(K0K1)) / s^2, written straight into p (r0 = p0 = b,
  // x0 = 0), with the per-column partials of <b, b> in the <v, v> slot
  const uint32_t cols = pl->cfg.image1;
  TSN_CHECK(op_load_samples(pl, measurement, grid_norm(pl), s));
  TSN_CHECK(op_adjoint_grid(pl, s));
  TSN_CHECK(op_adjoint_image(pl, pl->p, nullptr, 0.0f, inv_scale_sq, pl->partial, s, pl->p));
  // rho_k lives in pl->rho[k]; the final kernel writes sqrt(rho_k) to history.  The first forward FFT pass
  // writes rho_0 = <b, b> and the status.  Iteration k > 0: the first forward
  // pass applies step k - 1 (alpha from the partials of the last inverse pass:
  // x += alpha p, r -= alpha Ap, rho_k) and forms p_k = r_k + beta p_{k-1}
  // (r and p in ping-pong buffers; r_0 shares p_0's buffer, so the first
  // update writes r into a third one); the last inverse pass forms A p_k and
  // the partials of <p_k, A p_k>, Re<r_k, A p_k>, <A p_k, A p_k>.  One kernel
  // applies the last step and writes the result.  Six kernels per iteration.
  float2* p_cur = pl->p;
  float2* p_other = pl->p2;
  float2* r_cur = pl->p;
  float2* r_other = pl->r2;
  for (uint32_t it = 0; it < iterations; ++it) {
    FftArgs a = fft_args(pl);
    a.rho_hist = pl->rho;
    a.cg_status = pl->status;
    a.dot_count = cols;
    if (it == 0) {
      a.image_t = p_cur;
      a.rho0_dots = pl->partial + 2 * cols;
    } else {
      a.cg_r = r_cur;
      a.cg_r_new = r_other;
      a.cg_x = pl->x;
      a.x_zero = it == 1;
      a.cg_ap = pl->ap;
      a.cg_p_old = p_cur;
      a.cg_p_new = p_other;
      a.cg_dots = pl->partial;
      a.iter = it;
      std::swap(p_cur, p_other);
      float2* const r_done = r_cur;
      r_cur = r_other;
      r_other = it == 1 ? pl->r : r_done;  // r_0 was p_0's buffer, now p_other
    }
    TSN_CHECK(fft_forward(a, fft_out_kind(pl), s));
    TSN_CHECK(op_forward_samples(pl, nullptr, s));
    TSN_CHECK(op_adjoint_grid(pl, s));
    TSN_CHECK(op_adjoint_image(pl, pl->ap, p_cur, scaled_lambda, inv_scale_sq, pl->partial, s, r_cur));
  }
  // the last step, x_out = transpose(x) and the status copy in one kernel
  hipLaunchKernelGGL(cg_final_out_kernel, dim3(div_up(pl->cfg.image0, 32), div_up(pl->cfg.image1, 32)),
                     dim3(kVecThreads), 0, s, pl->x, p_cur, pl->partial, cols, pl->rho, iterations - 1,
                     iterations == 1 ? 1 : 0, pl->status, status_out, pl->cfg.image1, pl->cfg.image0, x_out,
                     history);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

int check_cg_args(tsn_plan* pl, const void* measurement, uint32_t iterations, float operator_scale,
                  float lambda, const void* x_out, const void* residuals, const void* status) {
  if (!pl || !measurement || !x_out || !residuals || !status) return fail(TSN_ERR_ARGUMENT, "null argument");
  if (iterations == 0) return fail(TSN_ERR_ARGUMENT, "cg iterations must be positive");
  if (iterations >= kMaxCgIterations)
    return fail(TSN_ERR_ARGUMENT, "cg iterations must be below %u", kMaxCgIterations);
  if (!(operator_scale > 0.0f) || !std::isfinite(operator_scale))
    return fail(TSN_ERR_ARGUMENT, "operator scale must be positive and finite");
  if (!(lambda >= 0.0f) || !std::isfinite(lambda)) return fail(TSN_ERR_ARGUMENT, "lambda must be >= 0");
  // rho_0 .. rho_iterations; a larger buffer replaces the old one, which stays
  // allocated until the plan is freed (a captured graph may still name it)
  if (iterations + 1 > pl->rho_cap) {
    TSN_CHECK(plan_alloc(pl, &pl->rho, iterations + 1));
    pl->rho_cap = iterations + 1;
  }
  return TSN_OK;
}


void destroy(tsn_plan* pl) {
  if (!pl) return;
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  for (void* p : pl->allocations) (void)hipFree(p);
  delete pl;
}

}  // namespace

extern "C" {

int tsn_plan_create(const tsn_plan_config* cfg, const tsn_direction* fwd, const tsn_direction* adj,
                    const void* smaps, const void* apodization, tsn_plan** out) {
  if (!cfg || !fwd || !adj || !smaps || !apodization || !out) return fail(TSN_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  const tsn_plan_config c = *cfg;
  if (c.coils == 0 || c.coils > static_cast<uint32_t>(kMaxCoils))
    return fail(TSN_ERR_ARGUMENT, "coils must be in 1..16");
  if (c.grid0 != 2 * c.image0 || c.grid1 != 2 * c.image1 || !fft_length_supported(c.grid0) ||
      !fft_length_supported(c.grid1))
    return fail(TSN_ERR_ARGUMENT, "grid must be 2x image with grid sides in {64, 128, 256, 512, 768}");
  if (c.precision < TSN_PREC_SPLIT16 || c.precision > TSN_PREC_SPLITBF16)
    return fail(TSN_ERR_ARGUMENT, "unknown precision %d", c.precision);
  const bool fp32 = c.precision == TSN_PREC_FP32;
  if (fwd->csr != fp32 || adj->csr != fp32)
    return fail(TSN_ERR_ARGUMENT, "FP32 plans need CSR directions, FP16 plans need packed directions");
  const uint32_t grid = c.grid0 * c.grid1;
  if (fwd->h.n_cols != grid || adj->h.output_size != grid || fwd->h.output_size != adj->h.n_cols)
    return fail(TSN_ERR_INCOMPATIBLE, "unexpected operator shape");
  if (c.precision == TSN_PREC_DENSE16) {
    TSN_CHECK(ensure_dense_tiles(fwd));
    TSN_CHECK(ensure_dense_tiles(adj));
  }
  if (!fp32) {
    const int want = c.precision == TSN_PREC_SPLITBF16 ? TSN_VALUES_BF16 : TSN_VALUES_FP16;
    if (fwd->values != want || adj->values != want)
      return fail(TSN_ERR_ARGUMENT, "precision %d needs %s packs", c.precision,
                  want == TSN_VALUES_BF16 ? "BF16" : "FP16");
  }
  auto* pl = new tsn_plan();
  pl->cfg = c;
  (void)hipGetDevice(&pl->device);
  pl->fwd = fwd;
  pl->adj = adj;
  pl->m = fwd->h.output_size;
  pl->blocks = div_up(c.coils, kCoilsPerBlock);
  pl->pixels = c.image0 * c.image1;
  pl->vec_blocks = div_up(pl->pixels, kVecThreads);
  const size_t px = pl->pixels;
  int rc = TSN_OK;
  auto A = [&](auto** p, size_t count) {
    if (rc == TSN_OK) rc = plan_alloc(pl, p, count);
  };
  A(&pl->tw0, c.grid0);
  A(&pl->tw1, c.grid1);
  A(&pl->smaps_t, px * c.coils);
  A(&pl->apod_t, px);
  const size_t grid_slots = fp32 ? c.coils : chunk_slots(c.coils);
  A(&pl->t_fwd, static_cast<size_t>(c.grid0) * c.image1 * chunk_slots(c.coils));
  A(&pl->t_inv, static_cast<size_t>(c.image1) * c.grid0 * chunk_slots(c.coils));
  A(&pl->grid, static_cast<size_t>(grid) * grid_slots);
  if (fp32) {
    A(&pl->pan32, static_cast<size_t>(grid) * c.coils);
    A(&pl->smp32, static_cast<size_t>(pl->m) * c.coils);
  } else {
    A(&pl->pan_il, static_cast<size_t>(pl->blocks) * grid * kN);
    A(&pl->smp_il, static_cast<size_t>(pl->blocks) * pl->m * kN);
    A(&pl->y_rows, static_cast<size_t>(pl->m) * c.coils);
  }
  for (float2** v : {&pl->x, &pl->r, &pl->r2, &pl->p, &pl->p2, &pl->ap, &pl->tmp}) A(v, px);
  A(&pl->partial, 3 * static_cast<size_t>(c.image1));
  // slots: kSlotFloats floats per packed chunk, kMaxCoils float2 per CSR chunk
  const size_t slot_floats = fp32 ? 2 * kMaxCoils : kSlotFloats;
  A(&pl->fwd_chunks.partials, static_cast<size_t>(fwd->n_slots) * slot_floats);
  A(&pl->fwd_chunks.counters, round_up(fwd->n_counters, 4u));
  A(&pl->adj_chunks.partials, static_cast<size_t>(adj->n_slots) * slot_floats);
  A(&pl->adj_chunks.counters, round_up(adj->n_counters, 4u));
  A(&pl->rho, kRhoInitial);
  pl->rho_cap = kRhoInitial;
  A(&pl->status, 1);
  if (rc == TSN_OK) {
    const auto t0 = twiddles(c.grid0), t1 = twiddles(c.grid1);
    const hipError_t e0 = hipMemcpy(pl->tw0, t0.data(), t0.size() * sizeof(float2), hipMemcpyHostToDevice);
    const hipError_t e1 = hipMemcpy(pl->tw1, t1.data(), t1.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e0 != hipSuccess || e1 != hipSuccess) rc = fail(TSN_ERR_RUNTIME, "twiddle upload failed");
  }
  if (rc == TSN_OK) {
    const uint64_t ns = static_cast<uint64_t>(c.coils) * px;
    hipLaunchKernelGGL(smaps_transpose_kernel, dim3(div_up(ns, 256)), dim3(256), 0, 0,
                       static_cast<const float2*>(smaps), c.coils, c.image0, c.image1, c.coils, pl->smaps_t);
    hipLaunchKernelGGL(apod_transpose_kernel, dim3(div_up(px, 256)), dim3(256), 0, 0,
                       static_cast<const float*>(apodization), c.image0, c.image1, pl->apod_t);
    // the adjoint grid's inactive rows are never written: zero it once here
    rc = zero_async(pl->grid, static_cast<size_t>(grid) * grid_slots * sizeof(float2), 0);
    // arrival counters start at zero (each launch's last arrivals reset them)
    for (const ChunkScratch* k : {&pl->fwd_chunks, &pl->adj_chunks})
      if (rc == TSN_OK && k->counters)
        rc = zero_async(k->counters, round_up(k == &pl->fwd_chunks ? fwd->n_counters : adj->n_counters, 4u) * 4, 0);
    if (rc == TSN_OK) {
      const hipError_t e = hipDeviceSynchronize();
      if (e != hipSuccess) rc = fail(TSN_ERR_RUNTIME, "plan setup: %s", hipGetErrorString(e));
    }
  }
  if (rc != TSN_OK) {
    destroy(pl);
    return rc;
  }
  *out = pl;
  return TSN_OK;
}

void tsn_plan_free(tsn_plan* pl) { destroy(pl); }

size_t tsn_plan_bytes(const tsn_plan* pl) { return pl ? pl->bytes : 0; }

int tsn_forward(tsn_plan* pl, const void* image, void* samples, void* stream) {
  if (!pl || !image || !samples) return fail(TSN_ERR_ARGUMENT, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  TSN_CHECK(transpose_image(static_cast<const float2*>(image), pl->cfg.image0, pl->cfg.image1, 1.0f, pl->tmp, s));
  TSN_CHECK(op_forward_grid(pl, pl->tmp, s));
  return op_forward_samples(pl, static_cast<float2*>(samples), s);
}

int tsn_adjoint(tsn_plan* pl, const void* samples, void* image, void* stream) {
  if (!pl || !image || !samples) return fail(TSN_ERR_ARGUMENT, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  TSN_CHECK(op_load_samples(pl, static_cast<const float2*>(samples), 1.0f, s));
  TSN_CHECK(op_adjoint_grid(pl, s));
  TSN_CHECK(op_adjoint_image(pl, pl->tmp, nullptr, 0.0f, 1.0f, nullptr, s));
  return transpose_image(pl->tmp, pl->cfg.image1, pl->cfg.image0, 1.0f, static_cast<float2*>(image), s);
}

int tsn_normal(tsn_plan* pl, const void* image, void* out, float inv_scale_sq, float scaled_lambda,
               void* stream) {
  if (!pl || !image || !out) return fail(TSN_ERR_ARGUMENT, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  TSN_CHECK(transpose_image(static_cast<const float2*>(image), pl->cfg.image0, pl->cfg.image1, 1.0f, pl->tmp, s));
  TSN_CHECK(op_normal(pl, pl->tmp, pl->ap, inv_scale_sq, scaled_lambda, nullptr, s));
  return transpose_image(pl->ap, pl->cfg.image1, pl->cfg.image0, 1.0f, static_cast<float2*>(out), s);
}

int tsn_data_consistency(tsn_plan* pl, const void* image, const void* measurement, const float* density,
                         float measurement_scale, void* out, void* stream) {
  if (!pl || !image || !measurement || !out) return fail(TSN_ERR_ARGUMENT, "null argument");
  if (!std::isfinite(measurement_scale)) return fail(TSN_ERR_ARGUMENT, "measurement scale must be finite");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t coils = pl->cfg.coils;
  DcTerms dc;
  dc.density = density;
  dc.scale = measurement_scale;
  TSN_CHECK(transpose_image(static_cast<const float2*>(image), pl->cfg.image0, pl->cfg.image1, 1.0f, pl->tmp, s));
  TSN_CHECK(op_forward_grid(pl, pl->tmp, s));
  if (is_fp32(pl)) {
    // CSR forward into the FP32 sample rows, then the residual in place (y coil-major)
    dc.y = static_cast<const float*>(measurement);
    dc.row_stride = 2;
    dc.coil_stride = 2 * pl->m;
    TSN_CHECK(launch_interp_csr(pl->fwd, pl->pan32, coils, 1, false, false, grid_norm(pl), pl->smp32, s,
                                &pl->fwd_chunks));
    TSN_CHECK(launch_dc_residual_rows(pl->smp32, pl->m, coils, dc, s));
  } else {
    // y to rows (one 64-byte read per sample row in the epilogue), residual fused
    const uint64_t n = static_cast<uint64_t>(coils) * pl->m;
    hipLaunchKernelGGL(samples_to_rows_kernel, dim3(div_up(n, 256)), dim3(256), 0, s,
                       static_cast<const float2*>(measurement), coils, pl->m, 1.0f, pl->y_rows);
    TSN_HIP(hipGetLastError());
    dc.y = reinterpret_cast<const float*>(pl->y_rows);
    dc.row_stride = 2 * coils;
    dc.coil_stride = 2;
    TSN_CHECK(launch_interp_dc(pl->fwd, interp_kind(pl), pl->pan_il, coils, grid_norm(pl), dc, pl->smp_il, s,
                               &pl->fwd_chunks));
  }
  TSN_CHECK(op_adjoint_grid(pl, s));
  TSN_CHECK(op_adjoint_image(pl, pl->ap, nullptr, 0.0f, 1.0f, nullptr, s));
  return transpose_image(pl->ap, pl->cfg.image1, pl->cfg.image0, 1.0f, static_cast<float2*>(out), s);
}

int tsn_cg(tsn_plan* pl, const void* measurement, uint32_t iterations, float operator_scale, float lambda,
           void* x_out, void* residuals_out, void* status_out, void* stream) {
  TSN_CHECK(check_cg_args(pl, measurement, iterations, operator_scale, lambda, x_out, residuals_out, status_out));
  return enqueue_cg(pl, static_cast<const float2*>(measurement), iterations, operator_scale, lambda,
                    static_cast<float2*>(x_out), static_cast<double*>(residuals_out),
                    static_cast<int*>(status_out), static_cast<hipStream_t>(stream));
}

int tsn_cg_graph_capture(tsn_plan* pl, const void* measurement, uint32_t iterations, float operator_scale,
                         float lambda, void* x_out, void* residuals_out, void* status_out, void* stream) {
  TSN_CHECK(check_cg_args(pl, measurement, iterations, operator_scale, lambda, x_out, residuals_out, status_out));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!s) return fail(TSN_ERR_ARGUMENT, "graph capture needs a non-null stream");
  if (pl->exec) {
    (void)hipGraphExecDestroy(pl->exec);
    pl->exec = nullptr;
  }
  if (pl->graph) {
    (void)hipGraphDestroy(pl->graph);
    pl->graph = nullptr;
  }
  TSN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
  const int rc = enqueue_cg(pl, static_cast<const float2*>(measurement), iterations, operator_scale, lambda,
                            static_cast<float2*>(x_out), static_cast<double*>(residuals_out),
                            static_cast<int*>(status_out), s);
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(s, &g);
  if (rc != TSN_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (end != hipSuccess) return fail(TSN_ERR_RUNTIME, "hipStreamEndCapture: %s", hipGetErrorString(end));
  pl->graph = g;
  TSN_HIP(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
  return TSN_OK;
}

int tsn_cg_graph_launch(tsn_plan* pl, void* stream) {
  if (!pl || !pl->exec) return fail(TSN_ERR_ARGUMENT, "no captured graph");
  TSN_HIP(hipGraphLaunch(pl->exec, static_cast<hipStream_t>(stream)));
  return TSN_OK;
}

}  // extern "C"
