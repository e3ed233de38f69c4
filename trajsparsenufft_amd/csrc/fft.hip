// In-house FFT endpoints for gfx950.
//
// One wave64 computes one length-K (K = 64*R, R in {1, 2, 4, 8, 12}) DFT: lane l holds
// elements l + 64 j; a radix-R pass in registers (zero-padded inputs skipped),
// then a 64-point DFT as two radix-8 passes through the wave's LDS slice.  A
// workgroup of eight waves transforms eight coils of one image column or grid
// row, so every global read and write is a contiguous block of rows with all
// coils interleaved.  Fused: SENSE expand x*S_c, apodization, zero-pad (forward
// first axis), FP16 hi/residual split into N16 panel rows holding hi | res << 16
// per channel (forward second axis), crop (inverse, pruned last radix-8 stage), SENSE
// combine sum_c conj(S_c) img_c, operator scaling, Tikhonov term and the CG dot
// partial <p, Ap> (inverse second axis).
//
// Forward model (reference nufft_custom_fft_forward_bench.cu:1600-1844,
// 538-747; semantics of torchkbnufft fft_and_scale): grid = DFT2(pad(S_c x s)),
// sign -1, unnormalised, no shift, grid row index k0*K1 + k1.  Inverse
// (reference :1463-1552): image = crop(IDFT2(grid)) unnormalised, scaled by the
// caller.
#include "fft_common.h"

namespace tsn {

// Stand-in source of the CG partials for passes that have none (read, never written).
__device__ const double kZeroPartials[1] = {0.0};

// Global-space view of a pointer chosen between sources: a generic (flat)
// load would count on the LDS wait counter too and force full waits.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const T* p) {
  return (const __attribute__((address_space(1))) T*)(p);
}

// ---------------------------------------------------------------- kernels

// Forward, first axis: image column n1 (all n0 < N0) of x*S_c*apod for eight
// coils -> T[k0][n1][c].  Image-space arrays are column-major ([n1][n0]).
template <int R>
__global__ __launch_bounds__(512) void fwd_axis0_kernel(FftArgs args) {
  const FftArgs a = problem_view(args, blockIdx.z);
  if (blockIdx.y * 8 >= a.coils) return;  // coil chunk beyond this problem (batched launches)
  extern __shared__ float2 lds[];
  __shared__ double red[24];
  constexpr int K = 64 * R;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // neighbouring columns on one XCD: their 64-byte rows of T fill whole L2 lines
  const uint32_t n1 = xcd_remap(blockIdx.x, gridDim.x), chunk = blockIdx.y;
  const uint32_t c = chunk * 8 + w;
  const bool active = c < a.coils;
  const WaveTwiddles<R> twp = load_twiddles<R>(a.tw0, lane);
  // issue every load first: the previous step's partials and rho (needed
  // first), then the image, or r, p, Ap and x, maps and apodization; the
  // reduction of the CG scalars then overlaps the latency of the later loads
  // (branch-free: selected sources, a zero array for the absent ones, so no
  // wait is forced where branches would merge; threads past dot_count load
  // its last entry and drop it below)
  const bool cg = a.cg_r != nullptr;
  const bool sums = cg || a.rho0_dots != nullptr;
  const uint32_t qi = sums ? min(threadIdx.x, a.dot_count - 1) : 0u;
  const uint32_t qs = cg ? a.dot_count : 0u;
  const double* q_src = cg ? a.cg_dots : (a.rho0_dots ? a.rho0_dots : kZeroPartials);
  const double* q_more = cg ? a.cg_dots : kZeroPartials;
  double q0 = gmem(q_src)[qi];
  double q1 = gmem(q_more)[qs + (cg ? qi : 0u)];
  double q2 = gmem(q_more)[2 * qs + (cg ? qi : 0u)];
  const double rho_old = *gmem(cg ? a.rho_hist + (a.iter - 1) : kZeroPartials);
  __builtin_amdgcn_sched_barrier(0);  // keep these loads ahead of the data loads
  const bool store_vec = chunk == 0 && w == 0;  // the wave that writes x, r and p_new
  // Loads are unconditional, at clamped indices (straight-line code lets every
  // load issue before the first wait; a load under a lane-dependent branch
  // would be waited for at the branch's end); out-of-range values are zeroed
  // afterwards.  Only j < JN can hold pixels: N0 = K/2.
  constexpr int JN = R >= 2 ? R / 2 : 1;
  const float2* src_r = a.cg_r ? a.cg_r : a.image_t;
  const float2* src_p = a.cg_r ? a.cg_p_old : a.image_t;
  const float2* src_a = a.cg_r ? a.cg_ap : a.image_t;
  const float2* src_s = a.smaps_t + (static_cast<size_t>(n1) * a.cstride + min(c, a.coils - 1)) * a.n0;
  const bool x_wave = a.cg_r && !a.x_zero && __builtin_amdgcn_readfirstlane(store_vec ? 1 : 0);
  float2 img[R], sm[R], pv[R], av[R], xv[R];
  float ap[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    img[j] = pv[j] = sm[j] = av[j] = xv[j] = make_float2(0.f, 0.f);
    ap[j] = 0.f;
    if (j >= JN) continue;
    const uint32_t n0 = min(lane + 64u * j, a.n0 - 1);
    const size_t pix = static_cast<size_t>(n1) * a.n0 + n0;
    img[j] = src_r[pix];
    pv[j] = src_p[pix];
    av[j] = src_a[pix];
    sm[j] = src_s[n0];
    ap[j] = a.apod_t[pix];
    if (x_wave) xv[j] = a.cg_x[pix];
  }
  // the partials' sums (every block, same order), while the data loads are in
  // flight: the previous step's <p, Ap>, Re<r, Ap>, <Ap, Ap>, or rho0 = <b, b>
  // (written by block (0, 0) with the status)
  if (sums) {
    if (threadIdx.x >= a.dot_count) q0 = q1 = q2 = 0.0;
    for (uint32_t i = threadIdx.x + 512; i < a.dot_count; i += 512) {
      q0 += gmem(q_src)[i];
      if (cg) {
        q1 += gmem(a.cg_dots)[a.dot_count + i];
        q2 += gmem(a.cg_dots)[2 * a.dot_count + i];
      }
    }
    block_sum3<8>(q0, q1, q2, red);
    if (!cg && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
      a.rho_hist[0] = q0;
      *a.cg_status = 1;
    }
  }
#pragma unroll
  for (int j = 0; j < JN; ++j) {
    if (!active || lane + 64u * j >= a.n0) {
      img[j] = pv[j] = sm[j] = av[j] = xv[j] = make_float2(0.f, 0.f);
      ap[j] = 0.f;
    }
  }
  if (cg) {
    const double pap = q0, rap = q1, apap = q2;
    const bool first_block = blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0;
    double alpha = 0.0;
    if (isfinite(pap) && pap > 0.0) alpha = rho_old / pap;
    else if (first_block) *a.cg_status = 0;
    const double rho_new = fmax(rho_old - 2.0 * alpha * rap + alpha * alpha * apap, 0.0);
    const double beta = rho_new / fmax(rho_old, 1e-30);
    if (first_block) a.rho_hist[a.iter] = rho_new;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const uint32_t n0 = lane + 64 * j;
      const float2 r_new = make_float2(static_cast<float>(img[j].x - alpha * av[j].x),
                                       static_cast<float>(img[j].y - alpha * av[j].y));
      img[j] = make_float2(static_cast<float>(r_new.x + beta * pv[j].x), static_cast<float>(r_new.y + beta * pv[j].y));
      if (active && n0 < a.n0 && store_vec) {
        const size_t pix = static_cast<size_t>(n1) * a.n0 + n0;
        a.cg_x[pix] = make_float2(static_cast<float>(xv[j].x + alpha * pv[j].x),
                                  static_cast<float>(xv[j].y + alpha * pv[j].y));
        a.cg_r_new[pix] = r_new;
        a.cg_p_new[pix] = img[j];
      }
    }
  }
  float2 x[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const float2 v = cmul(img[j], sm[j]);
    x[j] = make_float2(v.x * ap[j], v.y * ap[j]);
  }
  constexpr int TPL = (R * 8 + 63) / 64;
  float2 out[TPL][8];
  wave_fft<R, false, true, false>(x, lds + w * wave_lds_elems<R>(), a.tw0, twp, lane, out);
  __syncthreads();
  emit_exchange<R, false>(out, lds, lane, w);
  __syncthreads();
  const ChunkRows tl = chunk_rows(a.coils, static_cast<uint64_t>(a.k0) * a.n1);
  store_rows<K>(a.t_fwd + chunk * tl.cs + static_cast<size_t>(n1) * tl.rs, static_cast<uint64_t>(a.n1) * tl.rs, tl.rs,
                lds);
}

// The second-axis passes can take rows_per_block consecutive grid rows per
// block, the next row's loads issued as soon as the current row is in LDS.
// One row per block measured faster now that the transform synchronizes only
// its own wave (0.754 -> 0.745 ms per 512^2 / 8-coil reconstruction: two
// blocks per CU overlap one's transform with the other's memory phases), and
// at 768 (R = 12) the second row's registers pushed the kernels to 2-3 waves
// per SIMD.
template <int R>
constexpr uint32_t rows_per_block() { return 1; }

// Forward, second axis: grid row k0 -> panel rows k0*K1 + k1.  OUT 0: FP16
// hi/residual interleaved N16 panel blocks; OUT 1: float2 [row][coils];
// OUT 2: BF16 hi/residual interleaved panel blocks.
template <int R, int OUT>
__global__ __launch_bounds__(512) void fwd_axis1_kernel(FftArgs args) {
  const FftArgs a = problem_view(args, blockIdx.z);
  if (blockIdx.y * 8 >= a.coils) return;  // coil chunk beyond this problem (batched launches)
  const WaveTwiddles<R> twp = load_twiddles<R>(a.tw1, static_cast<int>(threadIdx.x & 63));
  extern __shared__ float2 lds[];
  constexpr int K = 64 * R;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr uint32_t kRowsPerBlock = rows_per_block<R>();
  const uint32_t row0 = blockIdx.x * kRowsPerBlock, chunk = blockIdx.y;
  const uint32_t cw = min(8u, a.coils - chunk * 8);
  const uint64_t grid_rows = static_cast<uint64_t>(a.k0) * K;
  // grid row k0 of T: N1 = K/2 rows (the plan checks grid = 2 x image)
  const ChunkRows tl = chunk_rows(a.coils, static_cast<uint64_t>(a.k0) * a.n1);
  auto t_row = [&](uint32_t k0) { return a.t_fwd + chunk * tl.cs + static_cast<size_t>(k0) * a.n1 * tl.rs; };
  ExchangeRegs<K / 2> rows;
  load_exchange<K / 2>(t_row(row0), tl.rs, cw, rows);
  for (uint32_t i = 0; i < kRowsPerBlock; ++i) {
    const uint32_t k0 = row0 + i;
    if (k0 >= a.k0) break;
    write_exchange<K / 2>(rows, lds);
    __syncthreads();
    if (i + 1 < kRowsPerBlock && k0 + 1 < a.k0) load_exchange<K / 2>(t_row(k0 + 1), tl.rs, cw, rows);
    float2 x[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const uint32_t n1 = lane + 64 * j;
      x[j] = n1 < a.n1 ? lds[xslot(n1, w)] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    constexpr int TPL = (R * 8 + 63) / 64;
    float2 out[TPL][8];
    wave_fft<R, false, true, false>(x, lds + w * wave_lds_elems<R>(), a.tw1, twp, lane, out);
    __syncthreads();
    emit_exchange<R, false>(out, lds, lane, w);
    __syncthreads();
    if constexpr (OUT == 0 || OUT == 2) {
      // two threads per panel row, one per coil half h (coils 4h .. 4h + 3):
      // the Re words 4h .. 4h + 3 and the Im words 8 + 4h .. 8 + 4h + 3 (hi |
      // res << 16) as two uint4 stores, each exchange element read once.
      // Thread e takes row (e >> 5) * 16 + (e & 15) and half (e >> 4) & 1, so
      // the 16 lanes of a group read one slot of 16 rows: distinct banks under
      // the exchange swizzle.
      uint4* base = reinterpret_cast<uint4*>(a.panel_il + (chunk * grid_rows + static_cast<uint64_t>(k0) * K) * 16);
#pragma unroll
      for (uint32_t e = threadIdx.x; e < static_cast<uint32_t>(K) * 2; e += 512u) {  // 512-thread blocks
        const uint32_t h = (e >> 4) & 1u, k1 = ((e >> 5) << 4) | (e & 15u);
        float2 v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) v[t] = lds[xslot(k1, 4 * h + t)];
        uint32_t re[4], im[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if constexpr (OUT == 2) {
            re[t] = split_bf16(v[t].x);
            im[t] = split_bf16(v[t].y);
          } else {
            split16_pair(v[t].x, v[t].y, re[t], im[t]);
          }
        }
        base[static_cast<size_t>(k1) * 4 + h] = make_uint4(re[0], re[1], re[2], re[3]);
        base[static_cast<size_t>(k1) * 4 + 2 + h] = make_uint4(im[0], im[1], im[2], im[3]);
      }
    } else {
      const ChunkRows gl = a.grid_blocked ? chunk_rows(a.coils, grid_rows) : ChunkRows{a.cstride, 8};
      store_rows<K>(a.panel32 + chunk * gl.cs + static_cast<uint64_t>(k0) * K * gl.rs, gl.rs,
                    a.grid_blocked ? gl.rs : cw, lds);
    }
    __syncthreads();  // the exchange area is refilled next
  }
}

// Inverse, second grid axis: grid row k0 ([k0*K1 + k1][coils]) -> IDFT along
// k1, keep n1 < N1 -> T2[n1][k0][c].  A block's two rows are neighbours, so
// its 64-byte rows of T2 fill whole 128-byte lines, and neighbouring pairs
// sit on one XCD.
template <int R>
__global__ __launch_bounds__(512) void inv_axis1_kernel(FftArgs args) {
  const FftArgs a = problem_view(args, blockIdx.z);
  if (blockIdx.y * 8 >= a.coils) return;  // coil chunk beyond this problem (batched launches)
  const WaveTwiddles<R> twp = load_twiddles<R>(a.tw1, static_cast<int>(threadIdx.x & 63));
  extern __shared__ float2 lds[];
  constexpr int K = 64 * R;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr uint32_t kRowsPerBlock = rows_per_block<R>();
  const uint32_t row0 = xcd_remap(blockIdx.x, gridDim.x) * kRowsPerBlock, chunk = blockIdx.y;
  const uint32_t cw = min(8u, a.coils - chunk * 8);
  const ChunkRows gl =
      a.grid_blocked ? chunk_rows(a.coils, static_cast<uint64_t>(a.k0) * K) : ChunkRows{a.cstride, 8};
  const ChunkRows tl = chunk_rows(a.coils, static_cast<uint64_t>(a.n1) * a.k0);
  auto g_row = [&](uint32_t k0) { return a.grid + chunk * gl.cs + static_cast<size_t>(k0) * K * gl.rs; };
  ExchangeRegs<K> rows;
  load_exchange<K>(g_row(row0), gl.rs, cw, rows);
  for (uint32_t i = 0; i < kRowsPerBlock; ++i) {
    const uint32_t k0 = row0 + i;
    if (k0 >= a.k0) break;
    write_exchange<K>(rows, lds);
    __syncthreads();
    if (i + 1 < kRowsPerBlock && k0 + 1 < a.k0) load_exchange<K>(g_row(k0 + 1), gl.rs, cw, rows);
    float2 x[R];
#pragma unroll
    for (int j = 0; j < R; ++j) x[j] = lds[xslot(lane + 64 * j, w)];
    __syncthreads();
    constexpr int TPL = (R * 8 + 63) / 64;
    float2 out[TPL][8];
    wave_fft<R, true, false, true>(x, lds + w * wave_lds_elems<R>(), a.tw1, twp, lane, out);
    __syncthreads();
    emit_exchange<R, true>(out, lds, lane, w);
    __syncthreads();
    store_rows<K / 2>(a.t_inv + chunk * tl.cs + static_cast<size_t>(k0) * tl.rs, static_cast<uint64_t>(a.k0) * tl.rs,
                      tl.rs, lds);
    __syncthreads();  // the exchange area is refilled next
  }
}

// Inverse, first grid axis for image column n1: IDFT along k0, keep n0 < N0,
// SENSE combine over all coils, * apod * out_scale, + lambda * reg, and the
// block's partial Re<reg, out> (CG denominator) in float64.
template <int R>
__global__ __launch_bounds__(512) void inv_axis0_kernel(FftArgs args) {
  const FftArgs a = problem_view(args, blockIdx.z);
  const WaveTwiddles<R> twp = load_twiddles<R>(a.tw0, static_cast<int>(threadIdx.x & 63));
  extern __shared__ float2 lds[];
  __shared__ double red[24];
  constexpr int K = 64 * R;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t n1 = blockIdx.x;
  constexpr int TPL = (R * 8 + 63) / 64;
  // the epilogue's operands are loaded first, so their latency overlaps the transforms
  // (at a clamped pixel, unconditionally: a load under a lane-dependent branch
  // would be waited for at the branch's end)
  const bool pixel = threadIdx.x < a.n0;
  const size_t pix = static_cast<size_t>(n1) * a.n0 + min(threadIdx.x, a.n0 - 1);
  // (an absent input reads the output column instead, unused)
  const float apv = a.apod_t[pix];
  const float2 regv = (a.reg_in ? a.reg_in : a.image_out)[pix];
  const float2 rv = (a.dot_r ? a.dot_r : a.image_out)[pix];
  float2 acc = make_float2(0.f, 0.f);
  const uint32_t chunks = (a.coils + 7) / 8;
  for (uint32_t chunk = 0; chunk < chunks; ++chunk) {
    const uint32_t cw = min(8u, a.coils - chunk * 8);
    // conj(S_c[n0]) of the wave's outputs, loaded with the chunk's data
    const uint32_t c = chunk * 8 + w;
    const float2* sm = a.smaps_t + (static_cast<size_t>(n1) * a.cstride + min(c, a.coils - 1)) * a.n0;
    float2 smv[TPL][4];
#pragma unroll
    for (int s = 0; s < TPL; ++s) {
      const int task = lane + 64 * s;
      const int k1 = task >> 3, k2a = task & 7;
#pragma unroll
      for (int k2b = 0; k2b < 4; ++k2b)
        smv[s][k2b] = task < R * 8 ? sm[k1 + R * (k2a + 8 * k2b)] : make_float2(0.f, 0.f);
    }
    const ChunkRows tl = chunk_rows(a.coils, static_cast<uint64_t>(a.n1) * a.k0);
    fill_exchange<K>(a.t_inv + chunk * tl.cs + static_cast<size_t>(n1) * K * tl.rs, tl.rs, cw, lds);
    __syncthreads();
    float2 x[R];
#pragma unroll
    for (int j = 0; j < R; ++j) x[j] = lds[xslot(lane + 64 * j, w)];
    __syncthreads();
    float2 out[TPL][8];
    wave_fft<R, true, false, true>(x, lds + w * wave_lds_elems<R>(), a.tw0, twp, lane, out);
    // multiply by conj(S_c[n0]) before the coil sum
#pragma unroll
    for (int s = 0; s < TPL; ++s) {
      if (lane + 64 * s >= R * 8) continue;
#pragma unroll
      for (int k2b = 0; k2b < 4; ++k2b)
        out[s][k2b] = (w < static_cast<int>(cw)) ? cmul_conj_b(out[s][k2b], smv[s][k2b]) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    emit_exchange<R, true>(out, lds, lane, w);
    __syncthreads();
    if (threadIdx.x < a.n0) {
#pragma unroll
      for (int i = 0; i < 8; ++i) {  // coil order
        const float2 v = lds[xslot(threadIdx.x, i)];
        acc.x += v.x;
        acc.y += v.y;
      }
    }
    __syncthreads();
  }
  double part = 0.0, part_r = 0.0, part_v = 0.0;
  if (pixel) {
    const float s = apv * a.out_scale;
    float2 v = make_float2(acc.x * s, acc.y * s);
    if (a.reg_in) {
      v.x += a.lambda * regv.x;
      v.y += a.lambda * regv.y;
      part = static_cast<double>(regv.x) * v.x + static_cast<double>(regv.y) * v.y;
    }
    if (a.dot_r) {
      part_r = static_cast<double>(rv.x) * v.x + static_cast<double>(rv.y) * v.y;
      part_v = static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y;
    }
    a.image_out[pix] = v;
  }
  if (a.dot_partial) {
    block_sum3<8>(part, part_r, part_v, red);
    if (threadIdx.x == 0) {
      a.dot_partial[n1] = part;
      if (a.dot_r) {
        a.dot_partial[a.n1 + n1] = part_r;
        a.dot_partial[2 * a.n1 + n1] = part_v;
      }
    }
  }
}

// ------------------------------------------------------------- launchers

template <int R>
static size_t lds_bytes_for() {
  return static_cast<size_t>(8) * wave_lds_elems<R>() * sizeof(float2);
}

static int radix_of(uint32_t k) {
  switch (k) {
    case 64: return 1;
    case 128: return 2;
    case 256: return 4;
    case 512: return 8;
    case 768: return 12;
    default: return 0;
  }
}

#define TSN_DISPATCH_R(Rv, EXPR) \
  switch (Rv) {                  \
    case 1: { constexpr int R = 1; EXPR; break; }   \
    case 2: { constexpr int R = 2; EXPR; break; }   \
    case 4: { constexpr int R = 4; EXPR; break; }   \
    case 8: { constexpr int R = 8; EXPR; break; }   \
    case 12: { constexpr int R = 12; EXPR; break; } \
    default: return fail(TSN_ERR_ARGUMENT, "unsupported FFT length"); \
  }

int fft_forward(const FftArgs& a, int out_kind, hipStream_t s, uint32_t problems) {
  const int r0 = radix_of(a.k0), r1 = radix_of(a.k1);
  const uint32_t chunks = (a.cstride + 7) / 8;
  TSN_DISPATCH_R(r0, hipLaunchKernelGGL(fwd_axis0_kernel<R>, dim3(a.n1, chunks, problems), dim3(512),
                                        lds_bytes_for<R>(), s, a));
  if (out_kind == 0) {
    TSN_DISPATCH_R(r1, hipLaunchKernelGGL((fwd_axis1_kernel<R, 0>), dim3(div_up(a.k0, rows_per_block<R>()), chunks, problems), dim3(512),
                                          lds_bytes_for<R>(), s, a));
  } else if (out_kind == 2) {
    TSN_DISPATCH_R(r1, hipLaunchKernelGGL((fwd_axis1_kernel<R, 2>), dim3(div_up(a.k0, rows_per_block<R>()), chunks, problems), dim3(512),
                                          lds_bytes_for<R>(), s, a));
  } else {
    TSN_DISPATCH_R(r1, hipLaunchKernelGGL((fwd_axis1_kernel<R, 1>), dim3(div_up(a.k0, rows_per_block<R>()), chunks, problems), dim3(512),
                                          lds_bytes_for<R>(), s, a));
  }
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

int fft_inverse(const FftArgs& a, hipStream_t s, uint32_t problems) {
  const int r0 = radix_of(a.k0), r1 = radix_of(a.k1);
  const uint32_t chunks = (a.cstride + 7) / 8;
  TSN_DISPATCH_R(r1, hipLaunchKernelGGL(inv_axis1_kernel<R>, dim3(div_up(a.k0, rows_per_block<R>()), chunks, problems), dim3(512),
                                        lds_bytes_for<R>(), s, a));
  TSN_DISPATCH_R(r0, hipLaunchKernelGGL(inv_axis0_kernel<R>, dim3(a.n1, 1, problems), dim3(512),
                                        lds_bytes_for<R>(), s, a));
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

bool fft_length_supported(uint32_t k) { return radix_of(k) != 0; }

}  // namespace tsn
