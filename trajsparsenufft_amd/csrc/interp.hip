// Interpolation on gfx950: 2:4 sparse MFMA (v_smfmac_f32_16x16x32_f16) over the
// reference's packed tiles, a dense-MFMA control, and the plain FP32 CSR path.
//
// Semantics follow the reference's fused complex kernel
// (experiments/production/src/nufft_spmma_bench.cu:377-663): for every 16-row
// group, sum over its K32 tiles of A_re/A_im (16x32, 2:4) times the panel rows
// named by the tile's 32 column ids (16 channels = Re/Im of 8 coils), complex
// combine (conjugate for the adjoint), scale, store or accumulate.
//
// MI355X design (not a translation): one wave64 per work item; the tile is
// re-laid out at upload so each lane fetches its MFMA operands with one 16-byte
// load (A_re|A_im) and one 2-byte load (both sparse index bytes); hi and
// residual panels accumulate into the same FP32 accumulators in one pass; the
// Re/Im combine is one cross-lane swap (lane n <-> n^8); long groups are split
// over several waves and reduced deterministically; block -> item mapping
// keeps Morton-adjacent groups on one XCD (shared L2).
#include <algorithm>
#include <numeric>

#include "tsn_internal.h"

using namespace tsn;

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

namespace tsn {

enum OutMode { kOutCoilMajor = 0, kOutRowMajor = 1, kOutSplit16 = 2 };
enum TileKind { kSparseSplit = 0, kDenseSplit = 1, kSparseHi = 2 };

// Block b runs on XCD b % 8; give each XCD a contiguous range of items.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t xcd = b & 7u, idx = b >> 3;
  const uint32_t per = nb >> 3, rem = nb & 7u;
  return xcd * per + min(xcd, rem) + idx;
}

struct InterpArgs {
  const WorkItem* items;
  uint32_t n_items;
  const WorkItem* long_items;
  uint32_t n_long;        // multiple of 8 (padding entries have group < 0)
  const int32_t* group_rows;
  const int32_t* tile_cols;
  const uint4* tile_a;
  const uint16_t* tile_idx;
  const uint4* tile_dense;
  const _Float16* panel_hi;
  const _Float16* panel_res;
  uint64_t panel_stride;  // halves between coil blocks of the input panel
  uint32_t coils;
  uint32_t n_out;         // output rows (for layouts)
  float scale;
  float* out_f32;         // float2 output (coil- or row-major)
  _Float16* out_hi;       // split output panel blocks
  _Float16* out_res;
};

template <int OM, bool CONJ, bool ACC>
__device__ __forceinline__ void interp_epilogue(const InterpArgs& a, int group, int b,
                                                const float4_t& acc_re, const float4_t& acc_im,
                                                int lane) {
  const int n = lane & 15, q = lane >> 4;
  const int4 rows = *reinterpret_cast<const int4*>(a.group_rows + group * 16 + q * 4);
  const int rr[4] = {rows.x, rows.y, rows.z, rows.w};
  const uint32_t coil = b * kCoilsPerBlock + (n & 7);
  const int part = n >> 3;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float other = __shfl_xor(acc_im[i], 8);
    float v;
    if (n < 8) v = CONJ ? acc_re[i] + other : acc_re[i] - other;
    else v = CONJ ? acc_re[i] - other : acc_re[i] + other;
    v *= a.scale;
    const int row = rr[i];
    if (row < 0) continue;
    if constexpr (OM == kOutSplit16) {
      const _Float16 hi = static_cast<_Float16>(v);
      const _Float16 res = static_cast<_Float16>(v - static_cast<float>(hi));
      const size_t o = (static_cast<size_t>(b) * a.n_out + row) * kN + n;
      a.out_hi[o] = hi;
      a.out_res[o] = res;
    } else {
      if (coil >= a.coils) continue;
      size_t o;
      if constexpr (OM == kOutCoilMajor) o = (static_cast<size_t>(coil) * a.n_out + row) * 2 + part;
      else o = (static_cast<size_t>(row) * a.coils + coil) * 2 + part;
      if constexpr (ACC) v += a.out_f32[o];
      a.out_f32[o] = v;
    }
  }
}

template <int KIND, int NB>
__device__ __forceinline__ void accumulate_tile(const InterpArgs& a, int4 c0, int4 c1, uint4 av,
                                                uint4 dv0, uint4 dv1, uint32_t w,
                                                float4_t (&acc_re)[NB], float4_t (&acc_im)[NB],
                                                int lane) {
  const int n = lane & 15;
  const int cols[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const _Float16* ph = a.panel_hi + b * a.panel_stride + n;
    half8_t bh;
#pragma unroll
    for (int j = 0; j < 8; ++j) bh[j] = ph[static_cast<size_t>(cols[j]) * kN];
    half8_t br;
    if constexpr (KIND != kSparseHi) {
      const _Float16* pr = a.panel_res + b * a.panel_stride + n;
#pragma unroll
      for (int j = 0; j < 8; ++j) br[j] = pr[static_cast<size_t>(cols[j]) * kN];
    }
    if constexpr (KIND == kDenseSplit) {
      const half8_t dre = __builtin_bit_cast(half8_t, dv0);
      const half8_t dim = __builtin_bit_cast(half8_t, dv1);
      acc_re[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(dre, bh, acc_re[b], 0, 0, 0);
      acc_im[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(dim, bh, acc_im[b], 0, 0, 0);
      acc_re[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(dre, br, acc_re[b], 0, 0, 0);
      acc_im[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(dim, br, acc_im[b], 0, 0, 0);
    } else {
      const half4_t are = __builtin_bit_cast(half4_t, make_uint2(av.x, av.y));
      const half4_t aim = __builtin_bit_cast(half4_t, make_uint2(av.z, av.w));
      const int idx = static_cast<int>(w);
      acc_re[b] = __builtin_amdgcn_smfmac_f32_16x16x32_f16(are, bh, acc_re[b], idx, 0, 0);
      acc_im[b] = __builtin_amdgcn_smfmac_f32_16x16x32_f16(aim, bh, acc_im[b], idx, 0, 1);
      if constexpr (KIND == kSparseSplit) {
        acc_re[b] = __builtin_amdgcn_smfmac_f32_16x16x32_f16(are, br, acc_re[b], idx, 0, 0);
        acc_im[b] = __builtin_amdgcn_smfmac_f32_16x16x32_f16(aim, br, acc_im[b], idx, 0, 1);
      }
    }
  }
}

// Accumulate tiles first, first + stride, ... < end, fetching the next tile's
// metadata while the current tile gathers.
template <int KIND, int NB>
__device__ __forceinline__ void run_tiles(const InterpArgs& a, int first, int end, int stride,
                                          float4_t (&acc_re)[NB], float4_t (&acc_im)[NB], int lane) {
  const int q = lane >> 4;
  int4 c0 = {0, 0, 0, 0}, c1 = {0, 0, 0, 0};
  uint4 av = {0, 0, 0, 0}, dv0 = {0, 0, 0, 0}, dv1 = {0, 0, 0, 0};
  uint32_t w = 0;
  auto fetch = [&](int t) {
    const int4* cp = reinterpret_cast<const int4*>(a.tile_cols + static_cast<size_t>(t) * kK + q * 8);
    c0 = cp[0];
    c1 = cp[1];
    if constexpr (KIND == kDenseSplit) {
      dv0 = a.tile_dense[(static_cast<size_t>(t) * 64 + lane) * 2];
      dv1 = a.tile_dense[(static_cast<size_t>(t) * 64 + lane) * 2 + 1];
    } else {
      av = a.tile_a[static_cast<size_t>(t) * 64 + lane];
      w = a.tile_idx[static_cast<size_t>(t) * 64 + lane];
    }
  };
  if (first < end) fetch(first);
  for (int t = first; t < end; t += stride) {
    const int4 k0 = c0, k1 = c1;
    const uint4 kav = av, kd0 = dv0, kd1 = dv1;
    const uint32_t kw = w;
    if (t + stride < end) fetch(t + stride);
    accumulate_tile<KIND, NB>(a, k0, k1, kav, kd0, kd1, kw, acc_re, acc_im, lane);
  }
}

// Blocks of kTileWaves waves.  The first n_long blocks each own one long group
// (waves stride over its tiles and are summed through LDS in wave order); the
// other blocks hold one ordinary group per wave, XCD-contiguous.
template <int KIND, int NB, int OM, bool CONJ, bool ACC>
__global__ __launch_bounds__(kTileWaves * 64) void interp_tiles_kernel(InterpArgs a) {
  __shared__ float4_t red[kTileWaves][NB][2][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float4_t acc_re[NB], acc_im[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    acc_re[b] = float4_t{0.f, 0.f, 0.f, 0.f};
    acc_im[b] = float4_t{0.f, 0.f, 0.f, 0.f};
  }
  if (blockIdx.x < a.n_long) {
    const WorkItem it = a.long_items[blockIdx.x];
    if (it.group < 0) return;  // padding block: uniform exit before the barrier
    run_tiles<KIND, NB>(a, it.tile_begin + wave, it.tile_end, kTileWaves, acc_re, acc_im, lane);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      red[wave][b][0][lane] = acc_re[b];
      red[wave][b][1][lane] = acc_im[b];
    }
    __syncthreads();
    if (wave != 0) return;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      float4_t re = red[0][b][0][lane], im = red[0][b][1][lane];
      for (int w2 = 1; w2 < kTileWaves; ++w2) {
        re += red[w2][b][0][lane];
        im += red[w2][b][1][lane];
      }
      interp_epilogue<OM, CONJ, ACC>(a, it.group, b, re, im, lane);
    }
    return;
  }
  const uint32_t rb = blockIdx.x - a.n_long;
  const uint32_t item_id = xcd_remap(rb, gridDim.x - a.n_long) * kTileWaves + wave;
  if (item_id >= a.n_items) return;
  const WorkItem it = a.items[item_id];
  run_tiles<KIND, NB>(a, it.tile_begin, it.tile_end, 1, acc_re, acc_im, lane);
#pragma unroll
  for (int b = 0; b < NB; ++b) interp_epilogue<OM, CONJ, ACC>(a, it.group, b, acc_re[b], acc_im[b], lane);
}

struct CsrArgs {
  const uint32_t* row_ptr;
  const int32_t* row_ids;
  const int32_t* cols;
  const float2* vals;
  const int32_t* short_rows;
  uint32_t n_short;
  const int32_t* long_rows;
  uint32_t n_long;  // multiple of 8 (padding entries are -1)
  const float2* in;
  uint32_t coils;
  uint32_t n_out;
  float scale;
  float2* out;
};

// Plain FP32 path: LPR lanes per CSR row (one coil each), FP32 complex
// coefficients, input rows float2 [row][coils].  Rows longer than
// kLongRowEntries get a whole block: 256/LPR lane groups stride over the
// entries and are summed in group order through LDS.
template <bool CONJ>
__device__ __forceinline__ void csr_accumulate(const CsrArgs& a, uint32_t k, uint32_t c, bool active,
                                               float& re, float& im) {
  const float2 v = a.vals[k];
  const float2 x = active ? a.in[static_cast<size_t>(a.cols[k]) * a.coils + c] : make_float2(0.f, 0.f);
  if (CONJ) {
    re += v.x * x.x + v.y * x.y;
    im += v.x * x.y - v.y * x.x;
  } else {
    re += v.x * x.x - v.y * x.y;
    im += v.x * x.y + v.y * x.x;
  }
}

template <int OM, bool ACC>
__device__ __forceinline__ void csr_store(const CsrArgs& a, int32_t row, uint32_t c, float re, float im) {
  float2 v = make_float2(re * a.scale, im * a.scale);
  size_t o;
  if constexpr (OM == kOutCoilMajor) o = static_cast<size_t>(c) * a.n_out + row;
  else o = static_cast<size_t>(row) * a.coils + c;
  if constexpr (ACC) {
    const float2 prior = a.out[o];
    v.x += prior.x;
    v.y += prior.y;
  }
  a.out[o] = v;
}

template <int LPR, int OM, bool CONJ, bool ACC>
__global__ __launch_bounds__(256) void interp_csr_kernel(CsrArgs a) {
  constexpr int kGroups = 256 / LPR;
  __shared__ float2 red[256];
  const uint32_t seg = threadIdx.x / LPR, c = threadIdx.x % LPR;
  const bool active = c < a.coils;
  if (blockIdx.x < a.n_long) {
    const int32_t task = a.long_rows[blockIdx.x];
    if (task < 0) return;  // padding block: uniform exit before the barrier
    float re0 = 0.f, im0 = 0.f;
    for (uint32_t k = a.row_ptr[task] + seg; k < a.row_ptr[task + 1]; k += kGroups)
      csr_accumulate<CONJ>(a, k, c, active, re0, im0);
    red[threadIdx.x] = make_float2(re0, im0);
    __syncthreads();
    if (seg != 0 || !active) return;
    float re = 0.f, im = 0.f;
    for (int g = 0; g < kGroups; ++g) {
      re += red[g * LPR + c].x;
      im += red[g * LPR + c].y;
    }
    csr_store<OM, ACC>(a, a.row_ids[task], c, re, im);
    return;
  }
  const uint32_t rb = blockIdx.x - a.n_long;
  const uint32_t idx = xcd_remap(rb, gridDim.x - a.n_long) * kGroups + seg;
  if (idx >= a.n_short) return;
  const int32_t task = a.short_rows[idx];
  const uint32_t begin = a.row_ptr[task], end = a.row_ptr[task + 1];
  float re0 = 0.f, im0 = 0.f, re1 = 0.f, im1 = 0.f;
  uint32_t k = begin;
  for (; k + 2 <= end; k += 2) {
    csr_accumulate<CONJ>(a, k, c, active, re0, im0);
    csr_accumulate<CONJ>(a, k + 1, c, active, re1, im1);
  }
  if (k < end) csr_accumulate<CONJ>(a, k, c, active, re0, im0);
  if (!active) return;
  csr_store<OM, ACC>(a, a.row_ids[task], c, re0 + re1, im0 + im1);
}

// hi = RN16(v), res = RN16(v - hi), v = in * scale; in float2 [coil][rows];
// output panel blocks [block][row][16].
__global__ __launch_bounds__(256) void split_panel_kernel(const float2* __restrict__ in, uint32_t rows,
                                                          uint32_t coils, uint32_t blocks, float scale,
                                                          _Float16* __restrict__ hi,
                                                          _Float16* __restrict__ res) {
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t total = static_cast<uint64_t>(blocks) * rows * kN;
  if (gtid >= total) return;
  const int n = static_cast<int>(gtid & 15);
  const uint64_t br = gtid >> 4;
  const uint32_t row = static_cast<uint32_t>(br % rows);
  const uint32_t b = static_cast<uint32_t>(br / rows);
  const uint32_t coil = b * kCoilsPerBlock + (n & 7);
  float v = 0.f;
  if (coil < coils) {
    const float2 x = in[static_cast<size_t>(coil) * rows + row];
    v = (n < 8 ? x.x : x.y) * scale;
  }
  const _Float16 h = static_cast<_Float16>(v);
  hi[gtid] = h;
  res[gtid] = static_cast<_Float16>(v - static_cast<float>(h));
}

// ------------------------------------------------------------- dispatch

template <int KIND, int NB, int OM>
static void launch_tiles_om(const InterpArgs& a, bool conj, bool acc, hipStream_t s) {
  const uint32_t blocks = a.n_long + div_up(a.n_items, kTileWaves);
#define TSN_LAUNCH(C, A)                                                                     \
  do {                                                                                       \
    if (blocks) hipLaunchKernelGGL((interp_tiles_kernel<KIND, NB, OM, C, A>), dim3(blocks),  \
                                   dim3(kTileWaves * 64), 0, s, a);                          \
  } while (0)
  if (conj) {
    if (acc) TSN_LAUNCH(true, true); else TSN_LAUNCH(true, false);
  } else {
    if (acc) TSN_LAUNCH(false, true); else TSN_LAUNCH(false, false);
  }
#undef TSN_LAUNCH
}

template <int KIND, int NB>
static void launch_tiles_nb(const InterpArgs& a, int om, bool conj, bool acc, hipStream_t s) {
  if (om == kOutCoilMajor) launch_tiles_om<KIND, NB, kOutCoilMajor>(a, conj, acc, s);
  else if (om == kOutRowMajor) launch_tiles_om<KIND, NB, kOutRowMajor>(a, conj, acc, s);
  else launch_tiles_om<KIND, NB, kOutSplit16>(a, conj, false, s);
}

template <int KIND>
static void launch_tiles(const InterpArgs& a, int om, bool conj, bool acc, hipStream_t s) {
  if (a.coils > kCoilsPerBlock) launch_tiles_nb<KIND, 2>(a, om, conj, acc, s);
  else launch_tiles_nb<KIND, 1>(a, om, conj, acc, s);
}

// Internal entry used by the plan and the C ABI.  All pointers are device
// pointers sized by the caller; panels hold ceil(coils/8) blocks.
int launch_interp(const tsn_direction* d, int kind, const void* hi, const void* res, uint32_t coils,
                  int out_mode, bool conj, bool accumulate, float scale, float* out_f32,
                  _Float16* out_hi, _Float16* out_res, hipStream_t stream) {
  if (!d || d->csr) return fail(TSN_ERR_ARGUMENT, "interpolation needs a packed (non-CSR) direction");
  if (coils == 0 || coils > static_cast<uint32_t>(kMaxCoils))
    return fail(TSN_ERR_ARGUMENT, "coils must be in 1..16 (got %u)", coils);
  if (!hi || (kind != kSparseHi && !res))
    return fail(TSN_ERR_ARGUMENT, "missing input panel");
  if (out_mode == kOutSplit16 ? (!out_hi || !out_res) : !out_f32)
    return fail(TSN_ERR_ARGUMENT, "missing output");
  if (kind == kDenseSplit && !d->tile_dense)
    return fail(TSN_ERR_ARGUMENT, "direction has no dense-control tiles");
  InterpArgs a{};
  a.items = d->items;
  a.n_items = d->n_items;
  a.long_items = d->long_items;
  a.n_long = d->n_long;
  a.group_rows = d->group_rows;
  a.tile_cols = d->tile_cols;
  a.tile_a = d->tile_a;
  a.tile_idx = d->tile_idx;
  a.tile_dense = d->tile_dense;
  a.panel_hi = static_cast<const _Float16*>(hi);
  a.panel_res = static_cast<const _Float16*>(res);
  a.panel_stride = static_cast<uint64_t>(d->h.n_cols) * kN;
  a.coils = coils;
  a.n_out = d->h.output_size;
  a.scale = scale;
  a.out_f32 = out_f32;
  a.out_hi = out_hi;
  a.out_res = out_res;
  if (kind == kSparseSplit) launch_tiles<kSparseSplit>(a, out_mode, conj, accumulate, stream);
  else if (kind == kDenseSplit) launch_tiles<kDenseSplit>(a, out_mode, conj, accumulate, stream);
  else launch_tiles<kSparseHi>(a, out_mode, conj, accumulate, stream);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

template <int LPR, int OM>
static void launch_csr_om(const tsn_direction* d, const float2* in, uint32_t coils, bool conj,
                          bool acc, float scale, float2* out, hipStream_t s) {
  CsrArgs a{d->row_ptr, d->row_ids, d->cols, d->vals, d->short_rows, d->n_short, d->long_rows,
            d->n_long_rows, in, coils, d->h.output_size, scale, out};
  const uint32_t blocks = d->n_long_rows + div_up(static_cast<uint64_t>(d->n_short) * LPR, 256);
#define TSN_CSR(C, A) \
  hipLaunchKernelGGL((interp_csr_kernel<LPR, OM, C, A>), dim3(blocks), dim3(256), 0, s, a)
  if (conj) {
    if (acc) TSN_CSR(true, true); else TSN_CSR(true, false);
  } else {
    if (acc) TSN_CSR(false, true); else TSN_CSR(false, false);
  }
#undef TSN_CSR
}

template <int LPR>
static void launch_csr_lpr(const tsn_direction* d, const float2* in, uint32_t coils, int om,
                           bool conj, bool acc, float scale, float2* out, hipStream_t s) {
  if (om == kOutCoilMajor) launch_csr_om<LPR, kOutCoilMajor>(d, in, coils, conj, acc, scale, out, s);
  else launch_csr_om<LPR, kOutRowMajor>(d, in, coils, conj, acc, scale, out, s);
}

int launch_interp_csr(const tsn_direction* d, const float2* in, uint32_t coils, int out_mode,
                      bool conj, bool accumulate, float scale, float2* out, hipStream_t stream) {
  if (!d || !d->csr) return fail(TSN_ERR_ARGUMENT, "FP32 interpolation needs a CSR direction");
  if (coils == 0 || coils > static_cast<uint32_t>(kMaxCoils))
    return fail(TSN_ERR_ARGUMENT, "coils must be in 1..16 (got %u)", coils);
  if (!in || !out) return fail(TSN_ERR_ARGUMENT, "null input/output");
  if (out_mode != kOutCoilMajor && out_mode != kOutRowMajor)
    return fail(TSN_ERR_ARGUMENT, "unsupported FP32 output layout %d", out_mode);
  if (coils <= 4) launch_csr_lpr<4>(d, in, coils, out_mode, conj, accumulate, scale, out, stream);
  else if (coils <= 8) launch_csr_lpr<8>(d, in, coils, out_mode, conj, accumulate, scale, out, stream);
  else launch_csr_lpr<16>(d, in, coils, out_mode, conj, accumulate, scale, out, stream);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

int launch_split_panel(const float2* in, uint32_t rows, uint32_t coils, float scale, _Float16* hi,
                       _Float16* res, hipStream_t stream) {
  if (!in || !hi || !res || rows == 0) return fail(TSN_ERR_ARGUMENT, "bad split arguments");
  if (coils == 0 || coils > static_cast<uint32_t>(kMaxCoils))
    return fail(TSN_ERR_ARGUMENT, "coils must be in 1..16 (got %u)", coils);
  const uint32_t blocks = div_up(coils, kCoilsPerBlock);
  const uint64_t total = static_cast<uint64_t>(blocks) * rows * kN;
  hipLaunchKernelGGL(split_panel_kernel, dim3(div_up(total, 256)), dim3(256), 0, stream, in, rows,
                     coils, blocks, scale, hi, res);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

// ------------------------------------------------------------- upload

template <typename T>
static int device_copy(const T* host, size_t count, T** out, size_t* bytes) {
  *out = nullptr;
  if (count == 0) return TSN_OK;
  TSN_HIP(hipMalloc(reinterpret_cast<void**>(out), count * sizeof(T)));
  TSN_HIP(hipMemcpy(*out, host, count * sizeof(T), hipMemcpyHostToDevice));
  *bytes += count * sizeof(T);
  return TSN_OK;
}

static void release(tsn_direction* d) {
  if (!d) return;
  void* ptrs[] = {d->group_offsets, d->group_rows, d->tile_cols, d->tile_a, d->tile_idx,
                  d->tile_dense, d->items, d->long_items, d->row_ptr, d->row_ids,
                  d->cols, d->vals, d->short_rows, d->long_rows};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete d;
}

static int upload_pack(const HostPack& p, tsn_direction** out) {
  TSN_CHECK(validate_pack(p));
  const tsn_pack_header& h = p.h;
  const size_t tiles = h.tiles;
  std::vector<uint4> tile_a(tiles * 64);
  std::vector<uint16_t> tile_idx(tiles * 64);
  std::vector<uint4> tile_dense(tiles * 128);
  for (size_t t = 0; t < tiles; ++t) {
    uint32_t rm_re[16], rm_im[16];
    decode_row_meta(p.real_meta.data() + t * 16, rm_re);
    decode_row_meta(p.imag_meta.data() + t * 16, rm_im);
    const uint16_t* cre = p.real_comp.data() + t * 256;
    const uint16_t* cim = p.imag_comp.data() + t * 256;
    for (int lane = 0; lane < 64; ++lane) {
      const int r = lane & 15, q = lane >> 4;
      const uint16_t* re4 = cre + r * 16 + 4 * q;
      const uint16_t* im4 = cim + r * 16 + 4 * q;
      uint4 v;
      v.x = re4[0] | (static_cast<uint32_t>(re4[1]) << 16);
      v.y = re4[2] | (static_cast<uint32_t>(re4[3]) << 16);
      v.z = im4[0] | (static_cast<uint32_t>(im4[1]) << 16);
      v.w = im4[2] | (static_cast<uint32_t>(im4[3]) << 16);
      tile_a[t * 64 + lane] = v;
      tile_idx[t * 64 + lane] = static_cast<uint16_t>(((rm_re[r] >> (8 * q)) & 0xFF) |
                                                      (((rm_im[r] >> (8 * q)) & 0xFF) << 8));
      // dense control: expanded row r, columns 8q .. 8q+7 (chunks 2q, 2q+1)
      uint16_t dre[8] = {0}, dim[8] = {0};
      for (int cc = 0; cc < 2; ++cc) {
        const int chunk = 2 * q + cc;
        const uint32_t code_re = (rm_re[r] >> (4 * chunk)) & 0xF;
        const uint32_t code_im = (rm_im[r] >> (4 * chunk)) & 0xF;
        dre[4 * cc + (code_re & 3)] = cre[r * 16 + 2 * chunk];
        dre[4 * cc + (code_re >> 2)] = cre[r * 16 + 2 * chunk + 1];
        dim[4 * cc + (code_im & 3)] = cim[r * 16 + 2 * chunk];
        dim[4 * cc + (code_im >> 2)] = cim[r * 16 + 2 * chunk + 1];
      }
      uint4 d0, d1;
      std::memcpy(&d0, dre, 16);
      std::memcpy(&d1, dim, 16);
      tile_dense[(t * 64 + lane) * 2] = d0;
      tile_dense[(t * 64 + lane) * 2 + 1] = d1;
    }
  }
  // ordinary groups: one wave each; long groups: one block each (padded to a
  // multiple of 8 blocks so the ordinary blocks keep their XCD mapping)
  std::vector<WorkItem> items, long_items;
  for (uint32_t g = 0; g < h.groups; ++g) {
    const int32_t b = static_cast<int32_t>(p.group_offsets[g]);
    const int32_t e = static_cast<int32_t>(p.group_offsets[g + 1]);
    (e - b <= kSplitAboveTiles ? items : long_items).push_back({static_cast<int32_t>(g), b, e, 0});
  }
  while (long_items.size() % 8) long_items.push_back({-1, 0, 0, 0});
  auto* d = new tsn_direction();
  d->h = h;
  (void)hipGetDevice(&d->device);
  d->n_items = static_cast<uint32_t>(items.size());
  d->n_long = static_cast<uint32_t>(long_items.size());
  size_t bytes = 0;
  int rc = TSN_OK;
  if (rc == TSN_OK) rc = device_copy(p.group_offsets.data(), p.group_offsets.size(), &d->group_offsets, &bytes);
  if (rc == TSN_OK) rc = device_copy(p.group_row_ids.data(), p.group_row_ids.size(), &d->group_rows, &bytes);
  if (rc == TSN_OK) rc = device_copy(p.tile_col_ids.data(), p.tile_col_ids.size(), &d->tile_cols, &bytes);
  if (rc == TSN_OK) rc = device_copy(tile_a.data(), tile_a.size(), &d->tile_a, &bytes);
  if (rc == TSN_OK) rc = device_copy(tile_idx.data(), tile_idx.size(), &d->tile_idx, &bytes);
  if (rc == TSN_OK) rc = device_copy(tile_dense.data(), tile_dense.size(), &d->tile_dense, &bytes);
  if (rc == TSN_OK) rc = device_copy(items.data(), items.size(), &d->items, &bytes);
  if (rc == TSN_OK) rc = device_copy(long_items.data(), long_items.size(), &d->long_items, &bytes);
  if (rc != TSN_OK) {
    release(d);
    return rc;
  }
  d->bytes = bytes;
  *out = d;
  return TSN_OK;
}

}  // namespace tsn

extern "C" {

int tsn_direction_upload(const tsn_host_pack* pack, tsn_direction** out) {
  if (!pack || !out) return fail(TSN_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  return upload_pack(*pack, out);
}

int tsn_direction_load(const char* real_dir, const char* imag_dir, tsn_direction** out) {
  if (!out) return fail(TSN_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  tsn_host_pack* p = nullptr;
  TSN_CHECK(tsn_host_pack_read(real_dir, imag_dir, &p));
  const int rc = upload_pack(*p, out);
  tsn_host_pack_free(p);
  return rc;
}

int tsn_direction_upload_csr(const tsn_kb_matrix* kb, int view, tsn_direction** out) {
  if (!kb || !out || (view != 0 && view != 1)) return fail(TSN_ERR_ARGUMENT, "bad argument");
  *out = nullptr;
  const uint32_t m = kb->rows, w = kb->width;
  const uint32_t grid = kb->grid0 * kb->grid1;
  std::vector<uint32_t> ptr;
  std::vector<int32_t> rows, cols;
  std::vector<float2> vals;
  tsn_pack_header h{};
  h.magic = kMagic;
  h.version = kVersion;
  h.n_channels = kN;
  h.m_tile = kM;
  h.k_tile = kK;
  h.nnz = m * w;
  if (view == 0) {
    ptr.resize(m + 1);
    for (uint32_t i = 0; i <= m; ++i) ptr[i] = i * w;
    rows.resize(m);
    std::iota(rows.begin(), rows.end(), 0);
    cols = kb->cols;
    vals.resize(static_cast<size_t>(m) * w);
    for (size_t k = 0; k < vals.size(); ++k) vals[k] = make_float2(kb->re[k], kb->im[k]);
    h.output_size = m;
    h.n_cols = grid;
    h.active_rows = m;
  } else {
    // exact transpose: rows = active grid points ascending, entries by sample ascending
    std::vector<uint32_t> count(grid + 1, 0);
    for (int32_t c : kb->cols) ++count[c + 1];
    std::vector<uint32_t> start(grid + 1, 0);
    for (uint32_t g = 0; g < grid; ++g) start[g + 1] = start[g] + count[g + 1];
    std::vector<int32_t> tcols(static_cast<size_t>(m) * w);
    std::vector<float2> tvals(tcols.size());
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (uint32_t i = 0; i < m; ++i)
      for (uint32_t j = 0; j < w; ++j) {
        const size_t k = static_cast<size_t>(i) * w + j;
        const uint32_t pos = fill[kb->cols[k]]++;
        tcols[pos] = static_cast<int32_t>(i);
        tvals[pos] = make_float2(kb->re[k], kb->im[k]);
      }
    ptr.push_back(0);
    for (uint32_t g = 0; g < grid; ++g) {
      if (count[g + 1] == 0) continue;
      rows.push_back(static_cast<int32_t>(g));
      ptr.push_back(start[g + 1]);
    }
    cols.swap(tcols);
    vals.swap(tvals);
    h.output_size = grid;
    h.n_cols = m;
    h.active_rows = static_cast<uint32_t>(rows.size());
  }
  std::vector<int32_t> short_rows, long_rows;
  for (uint32_t t = 0; t + 1 < ptr.size(); ++t)
    (ptr[t + 1] - ptr[t] <= static_cast<uint32_t>(kLongRowEntries) ? short_rows : long_rows)
        .push_back(static_cast<int32_t>(t));
  while (long_rows.size() % 8) long_rows.push_back(-1);
  auto* d = new tsn_direction();
  d->h = h;
  d->csr = true;
  d->n_short = static_cast<uint32_t>(short_rows.size());
  d->n_long_rows = static_cast<uint32_t>(long_rows.size());
  (void)hipGetDevice(&d->device);
  d->csr_rows = static_cast<uint32_t>(rows.size());
  size_t bytes = 0;
  int rc = device_copy(ptr.data(), ptr.size(), &d->row_ptr, &bytes);
  if (rc == TSN_OK) rc = device_copy(rows.data(), rows.size(), &d->row_ids, &bytes);
  if (rc == TSN_OK) rc = device_copy(cols.data(), cols.size(), &d->cols, &bytes);
  if (rc == TSN_OK) rc = device_copy(vals.data(), vals.size(), &d->vals, &bytes);
  if (rc == TSN_OK) rc = device_copy(short_rows.data(), short_rows.size(), &d->short_rows, &bytes);
  if (rc == TSN_OK) rc = device_copy(long_rows.data(), long_rows.size(), &d->long_rows, &bytes);
  if (rc != TSN_OK) {
    release(d);
    return rc;
  }
  d->bytes = bytes;
  *out = d;
  return TSN_OK;
}

int tsn_direction_header(const tsn_direction* d, tsn_pack_header* out) {
  if (!d || !out) return fail(TSN_ERR_ARGUMENT, "null argument");
  *out = d->h;
  return TSN_OK;
}

size_t tsn_direction_bytes(const tsn_direction* d) { return d ? d->bytes : 0; }

void tsn_direction_free(tsn_direction* d) { release(d); }

int tsn_interp(const tsn_direction* d, const void* hi, const void* res, void* out, uint32_t coils,
               int conjugate, int accumulate, float scale, int out_layout, void* stream) {
  if (out_layout != TSN_OUT_COIL_MAJOR && out_layout != TSN_OUT_ROW_MAJOR)
    return fail(TSN_ERR_ARGUMENT, "unknown output layout %d", out_layout);
  return launch_interp(d, res ? kSparseSplit : kSparseHi, hi, res, coils, out_layout, conjugate != 0,
                       accumulate != 0, scale, static_cast<float*>(out), nullptr, nullptr,
                       static_cast<hipStream_t>(stream));
}

int tsn_interp_dense(const tsn_direction* d, const void* hi, const void* res, void* out,
                     uint32_t coils, int conjugate, int accumulate, float scale, int out_layout,
                     void* stream) {
  if (out_layout != TSN_OUT_COIL_MAJOR && out_layout != TSN_OUT_ROW_MAJOR)
    return fail(TSN_ERR_ARGUMENT, "unknown output layout %d", out_layout);
  if (!res) return fail(TSN_ERR_ARGUMENT, "dense control needs hi and residual panels");
  return launch_interp(d, kDenseSplit, hi, res, coils, out_layout, conjugate != 0, accumulate != 0,
                       scale, static_cast<float*>(out), nullptr, nullptr,
                       static_cast<hipStream_t>(stream));
}

int tsn_interp_f32(const tsn_direction* d, const void* in, void* out, uint32_t coils, int conjugate,
                   int accumulate, float scale, int out_layout, void* stream) {
  return launch_interp_csr(d, static_cast<const float2*>(in), coils, out_layout, conjugate != 0,
                           accumulate != 0, scale, static_cast<float2*>(out),
                           static_cast<hipStream_t>(stream));
}

int tsn_split_panel(const void* in, uint32_t rows, uint32_t coils, float scale, void* hi, void* res,
                    void* stream) {
  return launch_split_panel(static_cast<const float2*>(in), rows, coils, scale,
                            static_cast<_Float16*>(hi), static_cast<_Float16*>(res),
                            static_cast<hipStream_t>(stream));
}

}  // extern "C"
