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
#include <cmath>
#include <mutex>
#include <numeric>

#include "tsn_internal.h"

using namespace tsn;

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

namespace tsn {

// kOutBlockRows (host side only): row-major float2 in blocks of 8 coils, block
// b of rows at b * n_out * 8 (64-byte rows), launched as kOutRowMajor with
// the block stride set.
enum OutMode { kOutCoilMajor = 0, kOutRowMajor = 1, kOutSplit16 = 2, kOutSplit16IL = 3, kOutSplitBF16IL = 4,
               kOutBlockRows = 5 };
// Tile kinds: FP16 tiles x FP16 hi + residual on smfmac, the same on dense
// MFMA, FP16 hi only, and BF16 tiles x BF16 hi + residual on smfmac.
enum TileKind { kSparseSplit = 0, kDenseSplit = 1, kSparseHi = 2, kSparseSplitBF16 = 3 };


struct InterpArgs {
  const WorkItem* items;
  uint32_t n_items;
  const ChunkItem* chunks;  // long-group chunks
  uint32_t n_chunks;
  const uint32_t* order;    // [n_items + n_chunks] launch order (kChunkTag marks chunks)
  float* partials;          // their partial slots [slot][kSlotFloats]
  uint32_t* counters;       // arrival counter per long group
  const int32_t* group_rows;
  const uint32_t* group_offsets;
  const uint32_t* tile_cols;  // column * 32 | last-tile flag (bit 31 of word 2 of each lane segment)
  uint32_t n_tiles, n_groups;
  const uint4* tile_a;
  const uint16_t* tile_idx;
  const uint4* tile_dense;
  const _Float16* panel_hi;
  const _Float16* panel_res;
  const uint32_t* panel_il;  // interleaved hi|res panel (plan-internal layout)
  uint64_t panel_stride;     // elements between coil blocks of the input panel
  uint32_t coils;
  uint32_t n_out;         // output rows (for layouts)
  float scale;
  float* out_f32;         // float2 output (coil- or row-major)
  _Float16* out_hi;       // split output panel blocks
  _Float16* out_res;
  uint32_t* out_il;       // interleaved split output panel blocks
  // row-major float2 output: (row, coil) at float2 (coil >> 3) * out_bs +
  // row * out_rs + (coil & 7) ([row][coils]: out_rs = coils, out_bs = 8)
  uint32_t out_rs;
  uint64_t out_bs;
  uint32_t coil_base;     // first coil of block 0 in the output (batched launches)
  uint32_t depth;         // tiles in flight ahead of the multiply: 0 = PipeDepth, else 1..3
  // Batched launches (one coil block of one problem per entry; NB = 1,
  // interleaved panels): items[i] belongs to problems[item_problem[i]].
  const struct InterpProblem* problems;
  const uint32_t* item_problem;
  // Data-consistency epilogue of split outputs: each value v becomes
  // (v - dc_scale * y) * w before the split.  y is read at
  // row * dc_row_stride + coil * dc_coil_stride + part (floats); null = off.
  const float* dc_meas;
  const float* dc_density;  // w per output row, or null (all ones)
  float dc_scale;
  uint32_t dc_row_stride, dc_coil_stride;
};

// One coil block of one problem of a batched interpolation launch.
struct InterpProblem {
  const int32_t* group_rows;
  const uint32_t* tile_cols;
  const uint4* tile_a;
  const uint16_t* tile_idx;
  const uint4* tile_dense;
  uint32_t n_tiles, n_groups, n_out, n_cols;
  uint32_t coils;       // coils of the problem (output row layouts)
  uint32_t coil_base;   // first coil of this block
  const uint32_t* panel_il;  // this block's input panel [n_cols][16]
  float* out_f32;       // the problem's float2 output
  uint32_t* out_il;     // this block's split output panel [n_out][16]
  uint32_t out_rs;      // row-major output strides (as InterpArgs)
  uint64_t out_bs;
};

// Global-space view of a plain pointer.  The kernel's loads and stores after
// the first tile go through it: a generic (flat) access counts on both the
// vector-memory and the LDS wait counters, and while one is outstanding the
// compiler drains every later wait to zero, which stalls the tile pipeline at
// each epilogue.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* glb(T* p) {
  return (__attribute__((address_space(1))) T*)(p);
}

__device__ __forceinline__ InterpArgs problem_args(const InterpArgs& a, uint32_t vp) {
  const InterpProblem& q = a.problems[vp];
  InterpArgs v = a;
  v.group_rows = q.group_rows;
  v.tile_cols = q.tile_cols;
  v.tile_a = q.tile_a;
  v.tile_idx = q.tile_idx;
  v.tile_dense = q.tile_dense;
  v.n_tiles = q.n_tiles;
  v.n_groups = q.n_groups;
  v.n_out = q.n_out;
  v.coils = q.coils;
  v.coil_base = q.coil_base;
  v.panel_il = q.panel_il;
  v.panel_stride = static_cast<uint64_t>(q.n_cols) * kN;  // block b's input panel follows at b strides
  v.out_f32 = q.out_f32;
  v.out_il = q.out_il;
  v.out_rs = q.out_rs;
  v.out_bs = q.out_bs;
  return v;
}

// hi | res << 16 of one value: FP16 (RN) or BF16 (RNE) halves, res = v - hi.
// v is pinned as an FP32 value first: otherwise the compiler may fold the
// multiply that produced it into the FP16 conversion (v_fma_mix, one rounding
// instead of two), and identical inputs would split differently in different
// kernels.
template <bool BF16>
__device__ __forceinline__ uint32_t split_word(float v) {
  asm volatile("" : "+v"(v));
  if constexpr (BF16) {
    const uint16_t hb = bf16_rne_bits(v);
    return static_cast<uint32_t>(hb) | (static_cast<uint32_t>(bf16_rne_bits(v - bf16_to_float(hb))) << 16);
  } else {
    const _Float16 h = static_cast<_Float16>(v);
    const _Float16 r = static_cast<_Float16>(v - static_cast<float>(h));
    return static_cast<uint32_t>(__builtin_bit_cast(uint16_t, h)) |
           (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, r)) << 16);
  }
}

// Data-consistency residual (prediction - scale * y) * w, every step rounded
// on its own (no contraction into fma), so that the fused epilogue and the
// materialized pass over stored predictions produce the same bits.
__device__ __forceinline__ float dc_residual(float pred, float y, float scale, float w) {
#pragma clang fp contract(off)
  const float m = y * scale;
  return (pred - m) * w;
}

__device__ __forceinline__ float dc_apply(const InterpArgs& a, float v, uint32_t coil, int row, int part) {
  asm volatile("" : "+v"(v));  // v is final here, as if loaded from memory
  const float y = coil < a.coils ? glb(a.dc_meas)[static_cast<size_t>(row) * a.dc_row_stride +
                                             static_cast<size_t>(coil) * a.dc_coil_stride + part]
                                 : 0.0f;
  const float w = a.dc_density ? glb(a.dc_density)[row] : 1.0f;
  return dc_residual(v, y, a.dc_scale, w);
}

template <int OM, bool CONJ, bool ACC>
__device__ __forceinline__ void interp_epilogue(const InterpArgs& a, const int4& rows, int b,
                                                const float4_t& acc_re, const float4_t& acc_im,
                                                int lane) {
  const int n = lane & 15;
  const int rr[4] = {rows.x, rows.y, rows.z, rows.w};
  const uint32_t coil = a.coil_base + b * kCoilsPerBlock + (n & 7);
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
    if constexpr (OM == kOutSplitBF16IL || OM == kOutSplit16IL) {
      if (a.dc_meas) v = dc_apply(a, v, coil, row, part);
      glb(a.out_il)[(static_cast<size_t>(b) * a.n_out + row) * kN + n] = split_word<OM == kOutSplitBF16IL>(v);
    } else if constexpr (OM == kOutSplit16) {
      asm volatile("" : "+v"(v));  // split the FP32 value (see split_word)
      const _Float16 hi = static_cast<_Float16>(v);
      const _Float16 res = static_cast<_Float16>(v - static_cast<float>(hi));
      const size_t o = (static_cast<size_t>(b) * a.n_out + row) * kN + n;
      glb(a.out_hi)[o] = hi;
      glb(a.out_res)[o] = res;
    } else {
      if (coil >= a.coils) continue;
      size_t o;
      if constexpr (OM == kOutCoilMajor) o = (static_cast<size_t>(coil) * a.n_out + row) * 2 + part;
      else o = ((coil >> 3) * a.out_bs + static_cast<size_t>(row) * a.out_rs + (coil & 7)) * 2 + part;
      if constexpr (ACC) v += glb(a.out_f32)[o];
      glb(a.out_f32)[o] = v;
    }
  }
}

// Buffer resources of one launch (wave-uniform, in scalar registers).  Every
// per-tile operand is a buffer load at a scalar tile offset plus a
// loop-invariant lane offset, and every gather one buffer load whose lane
// offset is one VALU op from the stored column word, so the tile loop spends
// almost no vector instructions on addresses.  Out-of-range offsets read 0.
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef __amdgpu_buffer_rsrc_t rsrc_t;

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, uint64_t bytes) {
  const uint32_t n = bytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<uint32_t>(bytes);
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), static_cast<short>(0), static_cast<int>(n),
                                           0x00020000);
}

template <int NB>
struct TileRsrc {
  rsrc_t cols, a, idx, dense, rows, hi[NB], res[NB];
};

template <int NB, bool IL>
__device__ __forceinline__ TileRsrc<NB> make_tile_rsrc(const InterpArgs& a) {
  TileRsrc<NB> r;
  const uint64_t tiles = a.n_tiles;
  r.cols = make_rsrc(a.tile_cols, tiles * kK * 4);
  r.a = make_rsrc(a.tile_a, tiles * 64 * 16);
  r.idx = make_rsrc(a.tile_idx, tiles * 64 * 2);
  r.dense = make_rsrc(a.tile_dense, a.tile_dense ? tiles * 64 * 32 : 0);
  r.rows = make_rsrc(a.group_rows, static_cast<uint64_t>(a.n_groups) * 64);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if constexpr (IL) {
      r.hi[b] = make_rsrc(a.panel_il + b * a.panel_stride, a.panel_stride * 4);
      r.res[b] = r.hi[b];
    } else {
      r.hi[b] = make_rsrc(a.panel_hi + b * a.panel_stride, a.panel_stride * 2);
      r.res[b] = make_rsrc(a.panel_res ? a.panel_res + b * a.panel_stride : a.panel_hi, a.panel_stride * 2);
    }
  }
  return r;
}

// Output rows of the lane's four accumulator rows (4q .. 4q+3) of group g.
template <int NB>
__device__ __forceinline__ int4 load_rows(const TileRsrc<NB>& r, int g, int lane) {
  const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(r.rows, (lane >> 4) * 16, g * 64, 0);
  return make_int4(static_cast<int>(v.x), static_cast<int>(v.y), static_cast<int>(v.z), static_cast<int>(v.w));
}

__device__ __forceinline__ int4 group_rows_of(const InterpArgs& a, uint32_t group, int lane) {
  const auto* p = glb(a.group_rows + static_cast<size_t>(group) * 16 + (lane >> 4) * 4);
  return make_int4(p[0], p[1], p[2], p[3]);
}

// Stage 1 of a tile: the lane's 8 device column words (tile columns
// 8q .. 8q+7).  A word holds the byte offset of the column's row in a half
// panel (column * 32); bit 31 of word 2 is set on the last tile of a group.
struct TileCols {
  u32x4_t c0, c1;
};

// The column words of a unit or chunk (at most kSplitAboveTiles tiles, 128
// bytes each) are copied once into the wave's LDS slice, so the gathers of a
// tile wait for an LDS read instead of a global load issued one step earlier
// (the chain column word -> gather address -> gather otherwise costs a full
// memory latency per tile).  Words past the range read neighbours or zeros
// (unused).
constexpr int kStageWords = kSplitAboveTiles * kK;  // per wave

template <int NB>
__device__ __forceinline__ void stage_cols(const TileRsrc<NB>& r, int first, int n, uint32_t* stage, int lane) {
  // both halves are loaded before either is stored (one memory round trip);
  // past the last tile the buffer range check returns zeros
  (void)n;
  first = __builtin_amdgcn_readfirstlane(first);
  const u32x4_t v0 = __builtin_amdgcn_raw_buffer_load_b128(r.cols, lane * 16, first * (kK * 4), 0);
  const u32x4_t v1 = __builtin_amdgcn_raw_buffer_load_b128(r.cols, 1024 + lane * 16, first * (kK * 4), 0);
  reinterpret_cast<u32x4_t*>(stage)[lane] = v0;
  reinterpret_cast<u32x4_t*>(stage)[64 + lane] = v1;
  __builtin_amdgcn_wave_barrier();
}

// Column words of the unit's k-th tile from the stage (the 16 lanes of a
// quarter read the same 32 bytes: LDS broadcast).
__device__ __forceinline__ TileCols staged_cols(const uint32_t* stage, int k, int lane) {
  const u32x4_t* p = reinterpret_cast<const u32x4_t*>(stage + k * kK) + (lane >> 4) * 2;
  TileCols c;
  c.c0 = p[0];
  c.c1 = p[1];
  return c;
}

// Stage 2 of a tile: A fragment, sparse index bytes, last-tile flag and the
// raw gathered panel words (kept unpacked until the multiply so no wait is
// forced before then).  IL: one dword (hi | res << 16) per row; otherwise one
// zero-extended half per row for hi (w[0..7]) and for the residual (w[8..15]).
template <int NB, bool IL>
struct TileOps {
  u32x4_t a0, a1;
  uint32_t idx, last;
  uint32_t w[NB][IL ? 8 : 16];
};

template <int KIND, int NB, bool IL>
__device__ __forceinline__ void fetch_ops(const TileRsrc<NB>& r, int t, const TileCols& c, int lane,
                                          TileOps<NB, IL>& o) {
  const uint32_t n = lane & 15;
  const uint32_t cw[8] = {c.c0.x, c.c0.y, c.c0.z, c.c0.w, c.c1.x, c.c1.y, c.c1.z, c.c1.w};
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if constexpr (IL) {
        // row byte offset in the interleaved panel: column * 64 + n * 4
        const uint32_t off = (cw[j] << 1) | (n << 2);
        o.w[b][j] = __builtin_amdgcn_raw_buffer_load_b32(r.hi[b], static_cast<int>(off), 0, 0);
      } else {
        const uint32_t off = (j == 2 ? (cw[j] & 0x7FFFFFFFu) : cw[j]) | (n << 1);
        o.w[b][j] = __builtin_amdgcn_raw_buffer_load_b16(r.hi[b], static_cast<int>(off), 0, 0);
        o.w[b][8 + j] = KIND == kSparseHi ? 0u
                                          : __builtin_amdgcn_raw_buffer_load_b16(r.res[b], static_cast<int>(off), 0, 0);
      }
    }
  }
  if constexpr (KIND == kDenseSplit) {
    o.a0 = __builtin_amdgcn_raw_buffer_load_b128(r.dense, lane * 32, t * (64 * 32), 0);
    o.a1 = __builtin_amdgcn_raw_buffer_load_b128(r.dense, lane * 32 + 16, t * (64 * 32), 0);
    o.idx = 0;
  } else {
    o.a0 = __builtin_amdgcn_raw_buffer_load_b128(r.a, lane * 16, t * (64 * 16), 0);
    o.a1 = u32x4_t{0u, 0u, 0u, 0u};
    o.idx = __builtin_amdgcn_raw_buffer_load_b16(r.idx, lane * 2, t * (64 * 2), 0);
  }
  o.last = c.c0.z >> 31;
}

// hi (sel 0x05040100) or residual (sel 0x07060302) halves of 8 words as half8.
__device__ __forceinline__ half8_t pack_halves(const uint32_t* w, uint32_t sel) {
  u32x4_t v;
  v.x = __builtin_amdgcn_perm(w[1], w[0], sel);
  v.y = __builtin_amdgcn_perm(w[3], w[2], sel);
  v.z = __builtin_amdgcn_perm(w[5], w[4], sel);
  v.w = __builtin_amdgcn_perm(w[7], w[6], sel);
  return __builtin_bit_cast(half8_t, v);
}
constexpr uint32_t kSelLo = 0x05040100u, kSelHi = 0x07060302u;

template <int KIND, int NB, bool IL>
__device__ __forceinline__ void mma_tile(const TileOps<NB, IL>& o, float4_t (&acc_re)[NB],
                                         float4_t (&acc_im)[NB]) {
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const half8_t hi = pack_halves(o.w[b], kSelLo);
    half8_t res;
    if constexpr (IL) res = pack_halves(o.w[b], kSelHi);
    else res = pack_halves(o.w[b] + 8, kSelLo);
    if constexpr (KIND == kDenseSplit) {
      const half8_t dre = __builtin_bit_cast(half8_t, o.a0);
      const half8_t dim = __builtin_bit_cast(half8_t, o.a1);
      acc_re[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(dre, hi, acc_re[b], 0, 0, 0);
      acc_im[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(dim, hi, acc_im[b], 0, 0, 0);
      acc_re[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(dre, res, acc_re[b], 0, 0, 0);
      acc_im[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(dim, res, acc_im[b], 0, 0, 0);
    } else if constexpr (KIND == kSparseSplitBF16) {
      const bf16x4_t are = __builtin_bit_cast(bf16x4_t, make_uint2(o.a0.x, o.a0.y));
      const bf16x4_t aim = __builtin_bit_cast(bf16x4_t, make_uint2(o.a0.z, o.a0.w));
      const bf16x8_t bhi = __builtin_bit_cast(bf16x8_t, hi);
      const bf16x8_t bres = __builtin_bit_cast(bf16x8_t, res);
      const int idx = static_cast<int>(o.idx);
      acc_re[b] = __builtin_amdgcn_smfmac_f32_16x16x32_bf16(are, bhi, acc_re[b], idx, 0, 0);
      acc_im[b] = __builtin_amdgcn_smfmac_f32_16x16x32_bf16(aim, bhi, acc_im[b], idx, 0, 1);
      acc_re[b] = __builtin_amdgcn_smfmac_f32_16x16x32_bf16(are, bres, acc_re[b], idx, 0, 0);
      acc_im[b] = __builtin_amdgcn_smfmac_f32_16x16x32_bf16(aim, bres, acc_im[b], idx, 0, 1);
    } else {
      const half4_t are = __builtin_bit_cast(half4_t, make_uint2(o.a0.x, o.a0.y));
      const half4_t aim = __builtin_bit_cast(half4_t, make_uint2(o.a0.z, o.a0.w));
      const int idx = static_cast<int>(o.idx);
      acc_re[b] = __builtin_amdgcn_smfmac_f32_16x16x32_f16(are, hi, acc_re[b], idx, 0, 0);
      acc_im[b] = __builtin_amdgcn_smfmac_f32_16x16x32_f16(aim, hi, acc_im[b], idx, 0, 1);
      if constexpr (KIND == kSparseSplit) {
        acc_re[b] = __builtin_amdgcn_smfmac_f32_16x16x32_f16(are, res, acc_re[b], idx, 0, 0);
        acc_im[b] = __builtin_amdgcn_smfmac_f32_16x16x32_f16(aim, res, acc_im[b], idx, 0, 1);
      }
    }
  }
}

// Gather stages in flight ahead of the multiply: 2 for one coil block of the
// interleaved panel, 0 for two coil blocks of separate panels or of
// dense-control tiles, else 1 (deeper stages measured no faster and cost
// occupancy).
template <int KIND, int NB, bool IL>
struct PipeDepth {
  static constexpr int value = (IL && NB == 1 && KIND != kDenseSplit) ? 2
                               : (NB == 2 && (!IL || KIND == kDenseSplit)) ? 0
                                                                          : 1;
};

// Stream tiles first .. end - 1 (at most kSplitAboveTiles, column words
// staged) of one wave through a software pipeline.  Step i reads the column
// words of tile i + D + 1 from the stage, issues the operand loads of tile
// i + D, then multiplies tile i (D: PipeDepth, or the autotuner's choice);
// operand loads return in issue order, so the multiply's wait leaves D tiles
// of loads in flight.
// Tile i lives in register slot i % U of both rings (U = max(D + 1, 2), the
// loop unrolled U times so every slot index is a constant).  Indices past the
// end are clamped to the last tile (valid, unused loads).  The tiles are
// those of the consecutive groups group, group + 1, ... (< group_end); with
// last_mask 1 the last tile of each runs the epilogue with the group's rows
// (loaded one group ahead) and resets the accumulators; with 0 (a chunk of a
// long group) everything accumulates.  One loop for both keeps the kernel's
// register count at that of the regular units.
template <int KIND, int NB, bool IL, int OM, bool CONJ, bool ACC, int D>
__device__ __forceinline__ void stream_tiles(const InterpArgs& a, const TileRsrc<NB>& r, int first, int end,
                                             int group, int group_end, uint32_t last_mask, float4_t (&acc_re)[NB],
                                             float4_t (&acc_im)[NB], int lane, const uint32_t* stage) {
  constexpr int U = D + 1 > 2 ? D + 1 : 2;
  // wave-uniform bounds: scalar control flow and scalar tile offsets
  first = __builtin_amdgcn_readfirstlane(first);
  end = __builtin_amdgcn_readfirstlane(end);
  int g = __builtin_amdgcn_readfirstlane(group);
  const int g_last = __builtin_amdgcn_readfirstlane(group_end) - 1;
  if (first >= end) return;
  const int n = end - first;
  const int last = end - 1;
  auto tile = [&](int i) { return min(first + i, last); };
  auto cols_of = [&](int i) { return staged_cols(stage, tile(i) - first, lane); };
  last_mask = __builtin_amdgcn_readfirstlane(last_mask);
  int4 rows = load_rows(r, g, lane);
  TileCols cols[U];
  TileOps<NB, IL> ops[U];
  // The scheduler must keep the issue order: the wait counter is in order, and
  // the loop's waits are computed over both the prologue and the back edge.
#pragma unroll
  for (int i = 0; i <= D; ++i) cols[i % U] = cols_of(i);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int i = 0; i < D; ++i) {
    fetch_ops<KIND, NB, IL>(r, tile(i), cols[i % U], lane, ops[i % U]);
    __builtin_amdgcn_sched_barrier(0);
  }
  for (int base = 0; base < n; base += U) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int i = base + j;
      __builtin_amdgcn_sched_barrier(0);
      // loads run unconditionally (clamped), only the multiply is skipped past
      // the end, so no loaded register is merged across a branch
      cols[(j + D + 1) % U] = cols_of(i + D + 1);
      __builtin_amdgcn_sched_barrier(0);
      fetch_ops<KIND, NB, IL>(r, tile(i + D), cols[(j + D) % U], lane, ops[(j + D) % U]);
      if (i < n) {
        mma_tile<KIND, NB, IL>(ops[j], acc_re, acc_im);
        if (__builtin_amdgcn_readfirstlane(ops[j].last) & last_mask) {
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            interp_epilogue<OM, CONJ, ACC>(a, rows, b, acc_re[b], acc_im[b], lane);
            acc_re[b] = float4_t{0.f, 0.f, 0.f, 0.f};
            acc_im[b] = float4_t{0.f, 0.f, 0.f, 0.f};
          }
          g = min(g + 1, g_last);
          rows = load_rows(r, g, lane);
        }
      }
    }
  }
}

// After a chunk of a long group has streamed its tiles: publish the partial
// sums, and if this wave is the group's last arrival, add every chunk's slot in
// chunk order and write the group's rows.  Slots are written and read with
// device-scope (sc1) 16-byte buffer accesses, the access a relaxed agent-scope
// atomic of each word compiles to: they reach past the per-XCD L2 without a
// cache flush.  The stores complete (vmcnt 0) before the arrival is counted,
// and the slots are read only after the count shows every chunk done.  A full
// agent-scope fence here would write back the whole L2 per chunk (measured:
// the launch went from 29 to 230 us).  The last wave resets the counter for
// the next launch.  No wave waits on another.
// The protocol rests on gfx950's mapping of device-scope accesses (sc1
// stores and loads reach the coherent level past the per-XCD L2, in issue
// order with respect to the counter's atomic after the store wait) rather
// than on language-level acquire/release; other targets are refused here.
#if defined(__HIP_DEVICE_COMPILE__) && !defined(__gfx950__)
#error "the chunk-slot protocol of the interpolation kernels is written for gfx950"
#endif
constexpr int kCpolDevice = 16;  // cache policy bits of a buffer access: sc1 (device scope)
constexpr uint32_t kSlotBatch = 8;  // slots loaded before the first add (one round trip for <= 8 chunks)
// Slot byte offsets are 32-bit buffer offsets: at most 2^31 bytes of slots.
constexpr uint32_t kMaxSlots = (1u << 31) / (kSlotFloats * 4);

// WB > 1 (one wave per coil block of the entry): wave wb publishes its
// accumulators as block wb, the group's last arrival of all chunks and waves
// adds the slots of every block; a holds the unshifted (block 0) arguments.
template <int OM, bool CONJ, bool ACC, int NB, int WB = 1>
__device__ __forceinline__ void chunk_finish(const InterpArgs& a, const ChunkItem& c, float4_t (&acc_re)[NB],
                                             float4_t (&acc_im)[NB], int lane, int wb = 0) {
  // slot layout: [coil block][re, im][lane][4] floats; slot offsets stay below
  // 2^32 bytes (kSlotFloats * 4 bytes per chunk)
  static_assert(NB * WB <= 2, "a slot holds two coil blocks");
  const rsrc_t slots = make_rsrc(a.partials, uint64_t{1} << 32);
  const int mine = static_cast<int>((c.slot0 + c.chunk) * kSlotFloats * 4);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, acc_re[b]), slots,
                                           (((b + wb) * 2 + 0) * 64 + lane) * 16, mine, kCpolDevice);
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, acc_im[b]), slots,
                                           (((b + wb) * 2 + 1) * 64 + lane) * 16, mine, kCpolDevice);
  }
  __builtin_amdgcn_s_waitcnt(0);  // every slot store acknowledged
  asm volatile("" ::: "memory");
  uint32_t arrived = 0;
  if (lane == 0)
    arrived = __hip_atomic_fetch_add(a.counters + c.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  arrived = __builtin_amdgcn_readfirstlane(arrived);
  if (arrived + 1 != c.n_chunks * WB) return;
  asm volatile("" ::: "memory");
  if (lane == 0) __hip_atomic_store(a.counters + c.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int first = static_cast<int>(c.slot0 * kSlotFloats * 4);
  const int4 rows = group_rows_of(a, static_cast<uint32_t>(c.group), lane);
  const uint32_t n = c.n_chunks;
#pragma unroll
  for (int b = 0; b < NB * WB; ++b) {
    float4_t re = float4_t{0.f, 0.f, 0.f, 0.f}, im = re;
    // kSlotBatch slots are loaded (all before the first add), then added in
    // chunk order: ceil(n / kSlotBatch) memory round trips
    for (uint32_t q0 = 0; q0 < n; q0 += kSlotBatch) {
      float4_t pr[kSlotBatch], pi[kSlotBatch];
#pragma unroll
      for (uint32_t j = 0; j < kSlotBatch; ++j) {
        const int so = first + static_cast<int>(min(q0 + j, n - 1) * kSlotFloats * 4);
        pr[j] = __builtin_bit_cast(float4_t, __builtin_amdgcn_raw_buffer_load_b128(
                                                 slots, ((b * 2 + 0) * 64 + lane) * 16, so, kCpolDevice));
        pi[j] = __builtin_bit_cast(float4_t, __builtin_amdgcn_raw_buffer_load_b128(
                                                 slots, ((b * 2 + 1) * 64 + lane) * 16, so, kCpolDevice));
      }
#pragma unroll
      for (uint32_t j = 0; j < kSlotBatch; ++j) {
        if (q0 + j >= n) break;
        if (q0 + j == 0) {
          re = pr[j];
          im = pi[j];
        } else {
          re += pr[j];
          im += pi[j];
        }
      }
    }
    interp_epilogue<OM, CONJ, ACC>(a, rows, b, re, im, lane);
  }
}

// One wave per block, one entry of the launch order each (a regular unit, a
// run of whole groups, or a chunk of a long group), the entries of one XCD
// contiguous in pack order.  One-wave blocks let the dispatcher start a new
// entry whenever a wave slot frees, so uneven entries balance themselves
// across the CUs.  Units and chunks share one tile loop (one inlined copy
// keeps the register count of the regular units).
// WB = 2 (entries of two coil blocks, NB = 1): a block of two waves per
// entry, wave w computing coil block w.  Both stream the entry's tile operands
// (the second read comes from cache: the two waves share a CU), each with the
// one-block register budget and pipeline, and gather only its own panel.
// (the arguments by value: a reference to the kernel's by-value argument
// would make the compiler copy it to scratch and read it back through flat
// loads)
template <int KIND, int NB, bool IL, int OM, bool CONJ, bool ACC, int WB, int D>
__device__ __forceinline__ void tiles_entry(InterpArgs a) {
  static_assert(WB == 1 || NB == 1, "one wave per coil block covers one block");
  __shared__ uint32_t stages[WB][kStageWords];
  const int lane = threadIdx.x & 63;
  const int wb = WB > 1 ? static_cast<int>(threadIdx.x >> 6) : 0;
  uint32_t* stage = stages[wb];
  float4_t acc_re[NB], acc_im[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    acc_re[b] = float4_t{0.f, 0.f, 0.f, 0.f};
    acc_im[b] = float4_t{0.f, 0.f, 0.f, 0.f};
  }
  const uint32_t entry = glb(a.order)[xcd_remap(blockIdx.x, gridDim.x)];
  const bool is_chunk = (entry & kChunkTag) != 0;
  ChunkItem c{};
  WorkItem u;
  uint32_t problem = 0;
  if (is_chunk) {
    c = a.chunks[entry & ~kChunkTag];
    u = {c.group, c.group + 1, c.tile_begin, c.tile_end};
    problem = c.problem;
  } else {
    u = a.items[entry];
    if (a.problems) problem = glb(a.item_problem)[entry];
  }
  InterpArgs v = a;
  if constexpr (IL) {
    if (a.problems) v = problem_args(a, problem);
  }
  // this wave's coil block: input panel, output coils and split-panel block
  InterpArgs vw = v;
  if constexpr (WB > 1) {
    vw.panel_il = v.panel_il ? v.panel_il + wb * v.panel_stride : nullptr;
    vw.panel_hi = v.panel_hi ? v.panel_hi + wb * v.panel_stride : nullptr;
    vw.panel_res = v.panel_res ? v.panel_res + wb * v.panel_stride : nullptr;
    vw.coil_base = v.coil_base + wb * kCoilsPerBlock;
    const size_t out_block = static_cast<size_t>(wb) * v.n_out * kN;
    vw.out_il = v.out_il ? v.out_il + out_block : nullptr;
    vw.out_hi = v.out_hi ? v.out_hi + out_block : nullptr;
    vw.out_res = v.out_res ? v.out_res + out_block : nullptr;
  }
  if (u.tile_begin == u.tile_end) {  // empty group: zero rows
    const int4 rows = group_rows_of(vw, static_cast<uint32_t>(u.group), lane);
#pragma unroll
    for (int b = 0; b < NB; ++b) interp_epilogue<OM, CONJ, ACC>(vw, rows, b, acc_re[b], acc_im[b], lane);
    return;
  }
  const TileRsrc<NB> r = make_tile_rsrc<NB, IL>(vw);
  stage_cols(r, u.tile_begin, u.tile_end - u.tile_begin, stage, lane);
  stream_tiles<KIND, NB, IL, OM, CONJ, ACC, D>(vw, r, u.tile_begin, u.tile_end, u.group, u.group_end,
                                               is_chunk ? 0u : 1u, acc_re, acc_im, lane, stage);
  if (is_chunk) chunk_finish<OM, CONJ, ACC, NB, WB>(v, c, acc_re, acc_im, lane, wb);
}

template <int KIND, int NB, bool IL, int OM, bool CONJ, bool ACC, int WB = 1>
__global__ __launch_bounds__(64 * WB) void interp_tiles_kernel(InterpArgs a) {
  tiles_entry<KIND, NB, IL, OM, CONJ, ACC, WB, PipeDepth<KIND, NB, IL>::value>(a);
}

// The same with another pipeline depth D (the autotuner's alternatives).
template <int KIND, int NB, bool IL, int OM, bool CONJ, bool ACC, int WB, int D>
__global__ __launch_bounds__(64 * WB) void interp_tiles_depth_kernel(InterpArgs a) {
  tiles_entry<KIND, NB, IL, OM, CONJ, ACC, WB, D>(a);
}


struct CsrArgs {
  const uint32_t* row_ptr;
  const int32_t* row_ids;
  const int32_t* cols;
  const float2* vals;
  const CsrItem* items;
  uint32_t n_items;
  float2* partials;    // [slots][kMaxCoils] partial sums of long-row chunks
  uint32_t* counters;  // one arrival counter per row of several chunks, zero between launches
  const float2* in;
  uint32_t n_cols;
  uint32_t coils;
  uint32_t n_out;
  float scale;
  float2* out;
  uint32_t skip_empty;  // the output's empty rows already hold zeros: runs of empty rows write nothing
};

// Plain FP32 path.  One wave per block takes one work item: a run of at most
// kCsrRowsPerItem consecutive short rows (each at most kCsrShortEntries
// entries), or one chunk of a longer row (at most kCsrChunkEntries entries).
// The item's column ids and FP32 complex coefficients are contiguous in the
// CSR arrays; the wave copies them into LDS with coalesced loads, all issued
// before the first wait.  LPR lanes (one coil each) then work on a row: the
// lane groups of a run take its rows, the lane groups of a chunk stride its
// entries and are summed in a fixed butterfly order.  Each group reads the
// column ids of a batch of entries from LDS (one broadcast address per group),
// issues all of the batch's gathers of input rows float2 [row][coils] and then
// accumulates in entry order, so a row's sum never depends on the schedule.
// A row of several chunks publishes one partial per chunk to a slot, and the
// wave that counts the row's last arrival adds the slots in chunk order and
// writes the row (the slot protocol of the packed kernel's long groups).
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef float float2_t __attribute__((ext_vector_type(2)));

template <bool CONJ>
__device__ __forceinline__ void cmac(float2_t v, float2_t x, bool valid, float& re, float& im) {
  float pr, pi;
  if (CONJ) {
    pr = v.x * x.x + v.y * x.y;
    pi = v.x * x.y - v.y * x.x;
  } else {
    pr = v.x * x.x - v.y * x.y;
    pi = v.x * x.y + v.y * x.x;
  }
  re = valid ? re + pr : re;
  im = valid ? im + pi : im;
}

// One batch: the gathers of entries e0, e0 + step, ... (B of them, those at or
// past end masked) are all issued before the first is used; the coefficients
// are read from LDS only when the products are formed, so a batch holds just
// its gathered rows in registers.
template <bool CONJ, int B>
__device__ __forceinline__ void csr_batch(const int32_t* s_col, const float2_t* s_val, uint32_t e0, uint32_t end,
                                          uint32_t step, rsrc_t in, uint32_t lane_off, uint32_t row_bytes,
                                          float& re, float& im) {
  float2_t x[B];
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const uint32_t e = e0 + i * step;
    const uint32_t col = static_cast<uint32_t>(s_col[e < end ? e : e0]);
    x[i] = __builtin_bit_cast(float2_t, __builtin_amdgcn_raw_buffer_load_b64(in, col * row_bytes + lane_off, 0, 0));
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const uint32_t e = e0 + i * step;
    cmac<CONJ>(s_val[e < end ? e : e0], x[i], e < end, re, im);
  }
}

// Entries first, first + step, ... < end of the LDS stage, in entry order:
// batches of 24 while more than 16 remain, then of 16 or 8 (a 36-entry row:
// 24 + 16, two memory round trips; 24 keeps the kernel at 5 waves per SIMD,
// where 32 needed 112 registers and 4 waves, measured no faster).
template <bool CONJ>
__device__ __forceinline__ void csr_row_sum(const int32_t* s_col, const float2_t* s_val, uint32_t first, uint32_t end,
                                            uint32_t step, rsrc_t in, uint32_t lane_off, uint32_t row_bytes,
                                            float& re, float& im) {
  uint32_t e0 = first;
  while (e0 < end) {
    const uint32_t left = (end - e0 + step - 1) / step;
    if (left > 16) {
      csr_batch<CONJ, 24>(s_col, s_val, e0, end, step, in, lane_off, row_bytes, re, im);
      e0 += 24 * step;
    } else if (left > 8) {
      csr_batch<CONJ, 16>(s_col, s_val, e0, end, step, in, lane_off, row_bytes, re, im);
      e0 += 16 * step;
    } else {
      csr_batch<CONJ, 8>(s_col, s_val, e0, end, step, in, lane_off, row_bytes, re, im);
      e0 += 8 * step;
    }
  }
}

template <int OM, bool ACC>
__device__ __forceinline__ void csr_store(const CsrArgs& a, int32_t row, uint32_t c, float re, float im) {
  float2_t v = float2_t{re * a.scale, im * a.scale};
  size_t o;
  if constexpr (OM == kOutCoilMajor) o = static_cast<size_t>(c) * a.n_out + row;
  else o = static_cast<size_t>(row) * a.coils + c;
  float2_t* out = reinterpret_cast<float2_t*>(a.out);
  if constexpr (ACC) v += glb(out)[o];
  glb(out)[o] = v;
}

template <int LPR, int OM, bool CONJ, bool ACC>
__global__ __launch_bounds__(64) void interp_csr_kernel(CsrArgs a) {
  constexpr int G = 64 / LPR;  // lane groups per wave
  __shared__ int32_t s_col[kCsrStageEntries];
  __shared__ float2_t s_val[kCsrStageEntries];
  const uint32_t lane = threadIdx.x;
  const CsrItem it = a.items[xcd_remap(blockIdx.x, gridDim.x)];
  const uint32_t n = it.e_end - it.e_begin;
  const uint32_t g = lane / LPR, c = lane % LPR;
  const bool active = c < a.coils;
  const bool chunk = it.n_chunks != 0;
  if (n == 0) {
    // a run of empty rows (at most kCsrEmptyRunRows): zeros, unless the
    // output holds them already or the launch accumulates; every row id is
    // loaded before the first store
    if (ACC || a.skip_empty) return;
    constexpr int kPer = (kCsrEmptyRunRows + G - 1) / G;
    int32_t orow[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) orow[k] = glb(a.row_ids)[min(it.row_begin + g + k * G, it.row_end - 1)];
#pragma unroll
    for (int k = 0; k < kPer; ++k)
      if (it.row_begin + g + k * G < it.row_end && active) csr_store<OM, false>(a, orow[k], c, 0.f, 0.f);
    return;
  }
  // the group's row of a run (its first; runs hold at most G rows when LPR <= 8)
  const uint32_t row = chunk ? it.row_begin : min(it.row_begin + g, it.row_end - 1);
  const uint32_t rb = glb(a.row_ptr)[row], re_ = glb(a.row_ptr)[row + 1];
  const int32_t out_row = glb(a.row_ids)[row];
  // stage the item's entries (clamped loads, every one issued before the first wait)
  int32_t cbuf[kCsrStageEntries / 64];
  float2_t vbuf[kCsrStageEntries / 64];
#pragma unroll
  for (int i = 0; i < kCsrStageEntries / 64; ++i) {
    const uint32_t e = it.e_begin + min(i * 64 + lane, n - 1);
    cbuf[i] = glb(a.cols)[e];
    vbuf[i] = glb(reinterpret_cast<const float2_t*>(a.vals))[e];
  }
#pragma unroll
  for (int i = 0; i < kCsrStageEntries / 64; ++i) {
    s_col[i * 64 + lane] = cbuf[i];
    s_val[i * 64 + lane] = vbuf[i];
  }
  __syncthreads();
  const uint32_t row_bytes = a.coils * 8;
  const rsrc_t in = make_rsrc(a.in, static_cast<uint64_t>(a.n_cols) * row_bytes);
  const uint32_t lane_off = (active ? c : 0) * 8;
  float re = 0.f, im = 0.f;
  if (!chunk) {
    uint32_t r = it.row_begin + g, b = rb, e = re_;
    int32_t orow = out_row;
    while (r < it.row_end) {
      re = 0.f;
      im = 0.f;
      csr_row_sum<CONJ>(s_col, s_val, b - it.e_begin, e - it.e_begin, 1, in, lane_off, row_bytes, re, im);
      if (active) csr_store<OM, ACC>(a, orow, c, re, im);
      r += G;
      if (r < it.row_end) {
        b = glb(a.row_ptr)[r];
        e = glb(a.row_ptr)[r + 1];
        orow = glb(a.row_ids)[r];
      }
    }
    return;
  }
  // one chunk of a long row: the groups stride the entries, then a butterfly over the groups
  csr_row_sum<CONJ>(s_col, s_val, g, n, G, in, lane_off, row_bytes, re, im);
#pragma unroll
  for (int off = LPR; off < 64; off <<= 1) {
    re += __shfl_xor(re, off);
    im += __shfl_xor(im, off);
  }
  if (it.n_chunks == 1) {
    if (g == 0 && active) csr_store<OM, ACC>(a, out_row, c, re, im);
    return;
  }
  const rsrc_t slots = make_rsrc(a.partials, uint64_t{1} << 32);
  if (g == 0 && active)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, float2_t{re, im}), slots, c * 8,
                                          static_cast<int>((it.slot0 + it.chunk) * kMaxCoils * 8), kCpolDevice);
  __builtin_amdgcn_s_waitcnt(0);  // the slot stores acknowledged
  asm volatile("" ::: "memory");
  uint32_t arrived = 0;
  if (lane == 0)
    arrived = __hip_atomic_fetch_add(a.counters + it.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  arrived = __builtin_amdgcn_readfirstlane(arrived);
  if (arrived + 1 != it.n_chunks) return;
  asm volatile("" ::: "memory");
  if (lane == 0) __hip_atomic_store(a.counters + it.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (g != 0 || !active) return;
  float sr = 0.f, si = 0.f;
  for (uint32_t q0 = 0; q0 < it.n_chunks; q0 += kSlotBatch) {
    float2_t p[kSlotBatch];
#pragma unroll
    for (uint32_t j = 0; j < kSlotBatch; ++j) {
      const uint32_t q = min(q0 + j, it.n_chunks - 1);
      p[j] = __builtin_bit_cast(float2_t, __builtin_amdgcn_raw_buffer_load_b64(
                                            slots, c * 8, static_cast<int>((it.slot0 + q) * kMaxCoils * 8),
                                            kCpolDevice));
    }
#pragma unroll
    for (uint32_t j = 0; j < kSlotBatch; ++j) {
      if (q0 + j >= it.n_chunks) break;
      sr += p[j].x;
      si += p[j].y;
    }
  }
  csr_store<OM, ACC>(a, out_row, c, sr, si);
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
  asm volatile("" : "+v"(v));  // split the FP32 value (see split_word)
  const _Float16 h = static_cast<_Float16>(v);
  hi[gtid] = h;
  res[gtid] = static_cast<_Float16>(v - static_cast<float>(h));
}

// Interleaved variant for the plan: one uint32 = hi | res << 16 per channel,
// FP16 (RN) or, with BF16, bfloat16 (RNE) halves.  A block converts 64 rows of
// one coil block: each wave reads 64 consecutive rows of one coil (512
// contiguous bytes), the tile is turned through LDS, and the panel rows go
// out as 64 contiguous bytes per row.
constexpr uint32_t kSplitRows = 64;

template <bool BF16>
__global__ __launch_bounds__(256) void split_panel_il_kernel(const float2* __restrict__ in, uint32_t rows,
                                                             uint32_t coils, float scale,
                                                             uint32_t* __restrict__ out) {
  __shared__ float2 tile[kSplitRows][kCoilsPerBlock + 1];
  const uint32_t row0 = blockIdx.x * kSplitRows, b = blockIdx.y;
#pragma unroll
  for (uint32_t k = 0; k < kSplitRows * kCoilsPerBlock / 256; ++k) {
    const uint32_t i = threadIdx.x + 256 * k;
    const uint32_t c = i / kSplitRows, r = i % kSplitRows;
    const uint32_t coil = b * kCoilsPerBlock + c, row = row0 + r;
    tile[r][c] = (coil < coils && row < rows) ? in[static_cast<size_t>(coil) * rows + row] : make_float2(0.f, 0.f);
  }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < kSplitRows * kN / 256; ++k) {
    const uint32_t w = threadIdx.x + 256 * k;
    const uint32_t r = w / kN, n = w % kN, row = row0 + r;
    if (row >= rows) continue;
    const float2 x = tile[r][n & 7];
    out[(static_cast<size_t>(b) * rows + row) * kN + n] = split_word<BF16>((n < 8 ? x.x : x.y) * scale);
  }
}

// Materialized data consistency: out = split((pred - scale * y) * w) with
// pred float2 [rows][coils] (a row-major interpolation output) and y read
// through the same strides as the fused epilogue.  One thread per panel word.
template <bool BF16>
__global__ __launch_bounds__(256) void dc_residual_kernel(const float* __restrict__ pred, const float* __restrict__ y,
                                                          const float* __restrict__ density, uint32_t rows,
                                                          uint32_t coils, uint32_t blocks, float scale,
                                                          uint32_t row_stride, uint32_t coil_stride,
                                                          uint32_t* __restrict__ out) {
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t total = static_cast<uint64_t>(blocks) * rows * kN;
  if (gtid >= total) return;
  const int n = static_cast<int>(gtid & 15);
  const uint64_t br = gtid >> 4;
  const uint32_t row = static_cast<uint32_t>(br % rows);
  const uint32_t b = static_cast<uint32_t>(br / rows);
  const uint32_t coil = b * kCoilsPerBlock + (n & 7);
  const int part = n >> 3;
  float p = 0.0f, m = 0.0f;
  if (coil < coils) {
    p = pred[(static_cast<size_t>(row) * coils + coil) * 2 + part];
    m = y[static_cast<size_t>(row) * row_stride + static_cast<size_t>(coil) * coil_stride + part];
  }
  const float w = density ? density[row] : 1.0f;
  out[gtid] = split_word<BF16>(dc_residual(p, m, scale, w));
}

// ------------------------------------------------------------- dispatch

// Pipeline depths the autotuner may pick besides PipeDepth: for the launches
// plans and batches make with interleaved panels (forward into a split
// panel, adjoint into float2 rows), one wave per entry: 1 or 3 tiles ahead
// for one coil block, 2 for two.
template <int KIND, int NB, bool IL, int OM, bool CONJ, bool ACC, int WB>
struct AltDepths {
  static constexpr bool kPlanLaunch = IL && KIND != kDenseSplit && !ACC && WB == 1 &&
                                      ((OM == kOutRowMajor && CONJ) ||
                                       ((OM == kOutSplit16IL || OM == kOutSplitBF16IL) && !CONJ));
};

template <int KIND, int NB, bool IL, int OM, bool CONJ, bool ACC, int WB>
static void launch_both(const InterpArgs& a, hipStream_t s) {
  const uint32_t blocks = a.n_chunks + a.n_items;
  if (!blocks) return;
  using Alt = AltDepths<KIND, NB, IL, OM, CONJ, ACC, WB>;
  if constexpr (Alt::kPlanLaunch) {
    if constexpr (NB == 1) {
      if (a.depth == 1) {
        hipLaunchKernelGGL((interp_tiles_depth_kernel<KIND, NB, IL, OM, CONJ, ACC, WB, 1>), dim3(blocks),
                           dim3(64 * WB), 0, s, a);
        return;
      }
      if (a.depth == 3) {
        hipLaunchKernelGGL((interp_tiles_depth_kernel<KIND, NB, IL, OM, CONJ, ACC, WB, 3>), dim3(blocks),
                           dim3(64 * WB), 0, s, a);
        return;
      }
    } else {
      if (a.depth == 2) {
        hipLaunchKernelGGL((interp_tiles_depth_kernel<KIND, NB, IL, OM, CONJ, ACC, WB, 2>), dim3(blocks),
                           dim3(64 * WB), 0, s, a);
        return;
      }
    }
  }
  hipLaunchKernelGGL((interp_tiles_kernel<KIND, NB, IL, OM, CONJ, ACC, WB>), dim3(blocks), dim3(64 * WB), 0, s, a);
}

template <int KIND, int NB, bool IL, int OM, int WB = 1>
static void launch_tiles_om(const InterpArgs& a, bool conj, bool acc, hipStream_t s) {
  constexpr bool kAccOk = OM != kOutSplit16 && OM != kOutSplit16IL && OM != kOutSplitBF16IL;
  if (conj) {
    if (kAccOk && acc) launch_both<KIND, NB, IL, OM, true, kAccOk, WB>(a, s);
    else launch_both<KIND, NB, IL, OM, true, false, WB>(a, s);
  } else {
    if (kAccOk && acc) launch_both<KIND, NB, IL, OM, false, kAccOk, WB>(a, s);
    else launch_both<KIND, NB, IL, OM, false, false, WB>(a, s);
  }
}

// Interleaved panels (plan-internal) pair with FP32 or interleaved split
// output; separate panels with FP32 or separate split output (checked by
// launch_interp).  WB = 2: two coil blocks, one wave each (NB = 1).
template <int KIND, int NB, bool IL, int WB = 1>
static void launch_tiles_nb(const InterpArgs& a, int om, bool conj, bool acc, hipStream_t s) {
  if constexpr (IL) {
    if (om == kOutCoilMajor) launch_tiles_om<KIND, NB, IL, kOutCoilMajor, WB>(a, conj, acc, s);
    else if (om == kOutRowMajor) launch_tiles_om<KIND, NB, IL, kOutRowMajor, WB>(a, conj, acc, s);
    else if (KIND == kSparseSplitBF16) launch_tiles_om<KIND, NB, IL, kOutSplitBF16IL, WB>(a, conj, false, s);
    else launch_tiles_om<KIND, NB, IL, kOutSplit16IL, WB>(a, conj, false, s);
  } else {
    if (om == kOutCoilMajor) launch_tiles_om<KIND, NB, IL, kOutCoilMajor>(a, conj, acc, s);
    else if (om == kOutRowMajor) launch_tiles_om<KIND, NB, IL, kOutRowMajor>(a, conj, acc, s);
    else launch_tiles_om<KIND, NB, IL, kOutSplit16>(a, conj, false, s);
  }
}

// Interleaved panels of two coil blocks: one wave for both blocks
// (coil_waves 1) or one wave per block (coil_waves 2).
template <int KIND>
static void launch_tiles_il2(const InterpArgs& a, int om, bool conj, bool acc, uint32_t coil_waves,
                             hipStream_t s) {
  if (coil_waves == 2) launch_tiles_nb<KIND, 1, true, 2>(a, om, conj, acc, s);
  else launch_tiles_nb<KIND, 2, true>(a, om, conj, acc, s);
}

template <int KIND>
static void launch_tiles(const InterpArgs& a, bool il, int om, bool conj, bool acc, uint32_t coil_waves,
                         hipStream_t s) {
  const bool two = a.coils > kCoilsPerBlock;
  if constexpr (KIND == kSparseSplitBF16) {  // interleaved panels only (checked by launch_interp)
    if (two) launch_tiles_il2<KIND>(a, om, conj, acc, coil_waves, s);
    else launch_tiles_nb<KIND, 1, true>(a, om, conj, acc, s);
    return;
  }
  if (il) {
    if (two) launch_tiles_il2<KIND>(a, om, conj, acc, coil_waves, s);
    else launch_tiles_nb<KIND, 1, true>(a, om, conj, acc, s);
  } else {
    if (two) launch_tiles_nb<KIND, 2, false>(a, om, conj, acc, s);
    else launch_tiles_nb<KIND, 1, false>(a, om, conj, acc, s);
  }
}

static int launch_interp_impl(const tsn_direction* d, int kind, const void* hi, const void* res, uint32_t coils,
                              int out_mode, bool conj, bool accumulate, float scale, float* out_f32,
                              _Float16* out_hi, _Float16* out_res, hipStream_t stream, bool interleaved,
                              uint32_t* out_il, const DcTerms* dc, const ChunkScratch* scratch) {
  if (!d || d->csr) return fail(TSN_ERR_ARGUMENT, "interpolation needs a packed (non-CSR) direction");
  if (coils == 0 || coils > static_cast<uint32_t>(kMaxCoils))
    return fail(TSN_ERR_ARGUMENT, "coils must be in 1..16 (got %u)", coils);
  if (!hi || (!interleaved && kind != kSparseHi && !res))
    return fail(TSN_ERR_ARGUMENT, "missing input panel");
  const bool block_rows = out_mode == kOutBlockRows;
  if (block_rows) out_mode = kOutRowMajor;
  const bool split_out = out_mode == kOutSplit16 || out_mode == kOutSplit16IL;
  if (split_out ? (interleaved ? !out_il : (!out_hi || !out_res)) : !out_f32)
    return fail(TSN_ERR_ARGUMENT, "missing output");
  if (interleaved ? out_mode == kOutSplit16 : out_mode == kOutSplit16IL)
    return fail(TSN_ERR_ARGUMENT, "output mode %d does not match the panel layout", out_mode);
  if (kind == kDenseSplit && !d->tile_dense)
    return fail(TSN_ERR_ARGUMENT, "direction has no dense-control tiles");
  if ((kind == kSparseSplitBF16) != (d->values == TSN_VALUES_BF16))
    return fail(TSN_ERR_ARGUMENT, "tile value format of the direction does not match the interpolation variant");
  if (kind == kSparseSplitBF16 && !interleaved)
    return fail(TSN_ERR_ARGUMENT, "BF16 tiles take an interleaved split panel");
  InterpArgs a{};
  a.items = d->items;
  a.n_items = d->n_items;
  a.chunks = d->chunks;
  a.n_chunks = d->n_chunks;
  a.order = d->order;
  a.partials = scratch ? scratch->partials : d->partials;
  a.counters = scratch ? scratch->counters : d->counters;
  if (d->n_chunks && (!a.partials || !a.counters)) return fail(TSN_ERR_ARGUMENT, "missing chunk scratch");
  a.group_rows = d->group_rows;
  a.group_offsets = d->group_offsets;
  a.tile_cols = d->tile_cols;
  a.n_tiles = d->h.tiles;
  a.n_groups = d->h.groups;
  a.tile_a = d->tile_a;
  a.tile_idx = d->tile_idx;
  a.tile_dense = d->tile_dense;
  a.panel_hi = static_cast<const _Float16*>(hi);
  a.panel_res = static_cast<const _Float16*>(res);
  a.panel_il = static_cast<const uint32_t*>(hi);
  a.panel_stride = static_cast<uint64_t>(d->h.n_cols) * kN;
  a.coils = coils;
  a.n_out = d->h.output_size;
  a.scale = scale;
  a.out_f32 = out_f32;
  a.out_hi = out_hi;
  a.out_res = out_res;
  a.out_il = out_il;
  a.out_rs = block_rows ? kCoilsPerBlock : coils;
  a.out_bs = block_rows ? static_cast<uint64_t>(a.n_out) * kCoilsPerBlock : kCoilsPerBlock;
  if (dc) {
    a.dc_meas = dc->y;
    a.dc_density = dc->density;
    a.dc_scale = dc->scale;
    a.dc_row_stride = dc->row_stride;
    a.dc_coil_stride = dc->coil_stride;
  }
  a.depth = d->pipe_depth;
  const uint32_t cw = d->coil_waves;
  if (kind == kSparseSplit) launch_tiles<kSparseSplit>(a, interleaved, out_mode, conj, accumulate, cw, stream);
  else if (kind == kDenseSplit) launch_tiles<kDenseSplit>(a, interleaved, out_mode, conj, accumulate, cw, stream);
  else if (kind == kSparseSplitBF16)
    launch_tiles<kSparseSplitBF16>(a, interleaved, out_mode, conj, accumulate, cw, stream);
  else launch_tiles<kSparseHi>(a, interleaved, out_mode, conj, accumulate, cw, stream);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

// Internal entry used by the plan and the C ABI.  All pointers are device
// pointers sized by the caller; panels hold ceil(coils/8) blocks.
int launch_interp(const tsn_direction* d, int kind, const void* hi, const void* res, uint32_t coils,
                  int out_mode, bool conj, bool accumulate, float scale, float* out_f32,
                  _Float16* out_hi, _Float16* out_res, hipStream_t stream, bool interleaved,
                  uint32_t* out_il, const ChunkScratch* scratch) {
  return launch_interp_impl(d, kind, hi, res, coils, out_mode, conj, accumulate, scale, out_f32, out_hi, out_res,
                            stream, interleaved, out_il, nullptr, scratch);
}

int launch_interp_dc(const tsn_direction* d, int kind, const uint32_t* panel, uint32_t coils, float scale,
                     const DcTerms& dc, uint32_t* out_il, hipStream_t s, const ChunkScratch* scratch) {
  if (!dc.y) return fail(TSN_ERR_ARGUMENT, "missing measurements");
  return launch_interp_impl(d, kind, panel, nullptr, coils, kOutSplit16IL, false, false, scale, nullptr, nullptr,
                            nullptr, s, true, out_il, &dc, scratch);
}

int launch_dc_residual(const float2* pred, uint32_t rows, uint32_t coils, const DcTerms& dc, bool bf16,
                       uint32_t* out_il, hipStream_t s) {
  if (!pred || !dc.y || !out_il) return fail(TSN_ERR_ARGUMENT, "null argument");
  if (coils == 0 || coils > static_cast<uint32_t>(kMaxCoils)) return fail(TSN_ERR_ARGUMENT, "coils must be in 1..16");
  const uint32_t blocks = div_up(coils, kCoilsPerBlock);
  const uint64_t total = static_cast<uint64_t>(blocks) * rows * kN;
  if (total == 0) return TSN_OK;
  const dim3 grid(static_cast<uint32_t>(div_up(total, 256)));
  const float* p = reinterpret_cast<const float*>(pred);
  if (bf16)
    hipLaunchKernelGGL(dc_residual_kernel<true>, grid, dim3(256), 0, s, p, dc.y, dc.density, rows, coils, blocks,
                       dc.scale, dc.row_stride, dc.coil_stride, out_il);
  else
    hipLaunchKernelGGL(dc_residual_kernel<false>, grid, dim3(256), 0, s, p, dc.y, dc.density, rows, coils, blocks,
                       dc.scale, dc.row_stride, dc.coil_stride, out_il);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

// FP32 rows float2 [rows][coils] in place: v <- (v - scale * y) * w.
__global__ __launch_bounds__(256) void dc_residual_rows_kernel(float* __restrict__ io, const float* __restrict__ y,
                                                               const float* __restrict__ density, uint32_t rows,
                                                               uint32_t coils, float scale, uint32_t row_stride,
                                                               uint32_t coil_stride) {
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t per_row = 2ull * coils;
  if (gtid >= per_row * rows) return;
  const uint32_t row = static_cast<uint32_t>(gtid / per_row);
  const uint32_t rem = static_cast<uint32_t>(gtid % per_row);
  const uint32_t coil = rem >> 1, part = rem & 1u;
  const float w = density ? density[row] : 1.0f;
  io[gtid] = dc_residual(io[gtid], y[static_cast<size_t>(row) * row_stride + static_cast<size_t>(coil) * coil_stride + part],
                         scale, w);
}

int launch_dc_residual_rows(float2* rows_io, uint32_t rows, uint32_t coils, const DcTerms& dc, hipStream_t s) {
  if (!rows_io || !dc.y) return fail(TSN_ERR_ARGUMENT, "null argument");
  const uint64_t total = 2ull * coils * rows;
  if (total == 0) return TSN_OK;
  hipLaunchKernelGGL(dc_residual_rows_kernel, dim3(static_cast<uint32_t>(div_up(total, 256))), dim3(256), 0, s,
                     reinterpret_cast<float*>(rows_io), dc.y, dc.density, rows, coils, dc.scale, dc.row_stride,
                     dc.coil_stride);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

template <int LPR, int OM>
static void launch_csr_om(const CsrArgs& a, bool conj, bool acc, hipStream_t s) {
#define TSN_CSR(C, A) \
  hipLaunchKernelGGL((interp_csr_kernel<LPR, OM, C, A>), dim3(a.n_items), dim3(64), 0, s, a)
  if (conj) {
    if (acc) TSN_CSR(true, true); else TSN_CSR(true, false);
  } else {
    if (acc) TSN_CSR(false, true); else TSN_CSR(false, false);
  }
#undef TSN_CSR
}

template <int LPR>
static void launch_csr_lpr(const CsrArgs& a, int om, bool conj, bool acc, hipStream_t s) {
  if (om == kOutCoilMajor) launch_csr_om<LPR, kOutCoilMajor>(a, conj, acc, s);
  else launch_csr_om<LPR, kOutRowMajor>(a, conj, acc, s);
}

int launch_interp_csr(const tsn_direction* d, const float2* in, uint32_t coils, int out_mode,
                      bool conj, bool accumulate, float scale, float2* out, hipStream_t stream,
                      const ChunkScratch* scratch, bool out_prezeroed) {
  if (!d || !d->csr) return fail(TSN_ERR_ARGUMENT, "FP32 interpolation needs a CSR direction");
  if (coils == 0 || coils > static_cast<uint32_t>(kMaxCoils))
    return fail(TSN_ERR_ARGUMENT, "coils must be in 1..16 (got %u)", coils);
  if (!in || !out) return fail(TSN_ERR_ARGUMENT, "null input/output");
  if (out_mode != kOutCoilMajor && out_mode != kOutRowMajor)
    return fail(TSN_ERR_ARGUMENT, "unsupported FP32 output layout %d", out_mode);
  // gathers address input rows with 32-bit byte offsets
  if (static_cast<uint64_t>(d->h.n_cols) * coils * 8 >= (uint64_t{1} << 32))
    return fail(TSN_ERR_ARGUMENT, "FP32 input of %u rows x %u coils exceeds 4 GiB", d->h.n_cols, coils);
  CsrArgs a{};
  a.row_ptr = d->row_ptr;
  a.row_ids = d->row_ids;
  a.cols = d->cols;
  a.vals = d->vals;
  a.items = d->csr_items;
  a.n_items = d->n_csr_items;
  a.partials = reinterpret_cast<float2*>(scratch ? scratch->partials : d->partials);
  a.counters = scratch ? scratch->counters : d->counters;
  if (d->n_slots && (!a.partials || !a.counters)) return fail(TSN_ERR_ARGUMENT, "missing chunk scratch");
  a.in = in;
  a.n_cols = d->h.n_cols;
  a.coils = coils;
  a.n_out = d->h.output_size;
  a.scale = scale;
  a.out = out;
  a.skip_empty = out_prezeroed ? 1u : 0u;
  if (out_prezeroed || accumulate) a.n_items = d->n_csr_active_items;  // the trailing empty runs write nothing
  if (a.n_items == 0) return TSN_OK;
  if (coils <= 4) launch_csr_lpr<4>(a, out_mode, conj, accumulate, stream);
  else if (coils <= 8) launch_csr_lpr<8>(a, out_mode, conj, accumulate, stream);
  else launch_csr_lpr<16>(a, out_mode, conj, accumulate, stream);
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

int launch_split_panel_il(const float2* in, uint32_t rows, uint32_t coils, float scale, uint32_t* out,
                          hipStream_t stream, int values) {
  if (!in || !out || rows == 0) return fail(TSN_ERR_ARGUMENT, "bad split arguments");
  if (coils == 0 || coils > static_cast<uint32_t>(kMaxCoils))
    return fail(TSN_ERR_ARGUMENT, "coils must be in 1..16 (got %u)", coils);
  if (values != TSN_VALUES_FP16 && values != TSN_VALUES_BF16)
    return fail(TSN_ERR_ARGUMENT, "unknown value format %d", values);
  const uint32_t blocks = div_up(coils, kCoilsPerBlock);
  const dim3 grid(div_up(rows, kSplitRows), blocks);
  if (values == TSN_VALUES_BF16)
    hipLaunchKernelGGL(split_panel_il_kernel<true>, grid, dim3(256), 0, stream, in, rows, coils, scale, out);
  else
    hipLaunchKernelGGL(split_panel_il_kernel<false>, grid, dim3(256), 0, stream, in, rows, coils, scale, out);
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
                  d->tile_dense, d->items, d->order, d->chunks, d->partials, d->counters, d->row_ptr, d->row_ids,
                  d->cols, d->vals, d->csr_items};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (void* p : d->retired) (void)hipFree(p);
  delete d;
}

// Longest entries first within each XCD's contiguous range of a launch order
// (block b runs on XCD b mod 8 and takes entry xcd_remap(b)), in windows of
// kLptWindow entries, about what one XCD holds resident: the waves dispatched
// last are then the shortest.  The sort is stable, so entries of equal length
// keep pack order.  Measured at 205,824 samples / 512^2 grid / 8 coils: G 18.5
// -> 17.9 us, G^T 19.7 -> 18.3 us.
constexpr uint32_t kLptWindow = 512;

template <typename F>
static void lpt_order(std::vector<uint32_t>& order, F tiles_of) {
  const uint32_t nb = static_cast<uint32_t>(order.size());
  for (uint32_t x = 0; x < 8; ++x) {
    const uint32_t lo = xcd_remap(x, nb), hi = x == 7 ? nb : xcd_remap(x + 1, nb);
    for (uint32_t b = lo; b < hi; b += kLptWindow)
      std::stable_sort(order.begin() + b, order.begin() + std::min(hi, b + kLptWindow),
                       [&](uint32_t p, uint32_t q) { return tiles_of(p) > tiles_of(q); });
  }
}

// Work schedule of a packed direction: the groups of at most
// kSplitAboveTiles tiles are cut, in pack order, into contiguous runs of whole
// groups holding at most unit_tiles tiles (a single group is never cut); an
// empty group is a unit of its own (zero epilogue).  Longer groups are
// chunks, fixed at upload.  unit_tiles 0 picks the default: the regular tiles
// over 12 waves per CU, clamped to 4..14 tiles.  The launch order then runs
// the longest entries of each XCD first (lpt_order).  Measured at 205,824
// samples / 512^2 grid / 8 coils: G 17.9 us with 14-tile units against 18.0
// (16) and 18.1 (12); G^T 18.3 us against 18.4 (16) and 20.1 (12); smaller
// units start more waves than the CUs hold at once.
static int schedule_units(tsn_direction* d, uint32_t unit_tiles) {
  const std::vector<uint32_t>& off = d->host_offsets;
  const uint32_t groups = d->h.groups;
  uint64_t regular_tiles = 0;
  for (uint32_t g = 0; g < groups; ++g) {
    const uint32_t n = off[g + 1] - off[g];
    if (n <= static_cast<uint32_t>(kSplitAboveTiles)) regular_tiles += n;
  }
  int64_t target = unit_tiles;
  if (target == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->device) != hipSuccess || cus <= 0)
      cus = 256;
    const uint64_t waves = static_cast<uint64_t>(cus) * 12;
    target = std::clamp<int64_t>(static_cast<int64_t>((regular_tiles + waves - 1) / waves), 4, 14);
  }
  // a unit's column words are staged in the wave's LDS slice: at most 16 tiles
  target = std::min<int64_t>(target, kSplitAboveTiles);
  std::vector<WorkItem> items;
  std::vector<uint32_t> order;
  uint32_t chunk = 0;  // chunks are numbered in group order (long_chunks)
  std::vector<int32_t> chunk_tiles;
  WorkItem cur{-1, -1, 0, 0};
  auto flush = [&]() {
    if (cur.group >= 0) {
      order.push_back(static_cast<uint32_t>(items.size()));
      items.push_back(cur);
    }
    cur = {-1, -1, 0, 0};
  };
  for (uint32_t g = 0; g < groups; ++g) {
    const int32_t b = static_cast<int32_t>(off[g]);
    const int32_t e = static_cast<int32_t>(off[g + 1]);
    const int32_t gi = static_cast<int32_t>(g);
    if (e - b > kSplitAboveTiles) {  // a chunked long group
      flush();
      const uint32_t n = static_cast<uint32_t>(e - b + kSplitAboveTiles - 1) / kSplitAboveTiles;
      for (uint32_t c = 0; c < n; ++c) {
        order.push_back(kChunkTag | chunk++);
        // a chunk ranks as if longer by its finishing chain, so chunks start first
        chunk_tiles.push_back(static_cast<int32_t>(chunk_begin(b, e, n, c + 1) - chunk_begin(b, e, n, c)) +
                              kChunkFinishTiles);
      }
      continue;
    }
    if (e == b) {
      flush();
      order.push_back(static_cast<uint32_t>(items.size()));
      items.push_back({gi, gi + 1, b, e});
      continue;
    }
    if (cur.group >= 0 && cur.tile_end - cur.tile_begin + (e - b) > target) flush();
    if (cur.group < 0) {
      cur = {gi, gi + 1, b, e};
    } else {
      cur.group_end = gi + 1;
      cur.tile_end = e;
    }
  }
  flush();
  lpt_order(order, [&](uint32_t e) {
    return (e & kChunkTag) ? chunk_tiles[e & ~kChunkTag] : items[e].tile_end - items[e].tile_begin;
  });
  WorkItem* new_items = nullptr;
  uint32_t* new_order = nullptr;
  size_t bytes = 0;
  int rc = device_copy(items.data(), items.size(), &new_items, &bytes);
  if (rc == TSN_OK) rc = device_copy(order.data(), order.size(), &new_order, &bytes);
  if (rc != TSN_OK) {
    if (new_items) (void)hipFree(new_items);
    return rc;
  }
  // Earlier schedules stay allocated until the direction is freed: a captured
  // graph may still hold their addresses in its kernel arguments.
  if (d->items) d->retired.push_back(d->items);
  if (d->order) d->retired.push_back(d->order);
  d->items = new_items;
  d->order = new_order;
  d->n_items = static_cast<uint32_t>(items.size());
  d->unit_tiles = static_cast<uint32_t>(target);
  d->host_items = items;
  d->host_order = order;
  d->bytes += bytes;
  return TSN_OK;
}

}  // namespace tsn

// Work items of a CSR direction in row order: runs of consecutive rows of at
// most kCsrShortEntries entries (at most kCsrRowsPerItem rows each), and each
// longer row cut into ceil(entries / kCsrChunkEntries) chunks of equal size
// (within one entry); a row of several chunks takes that many partial slots
// and one arrival counter.  Empty rows form runs of their own (at most
// kCsrEmptyRunRows rows) unless they follow a short row in a run.
std::vector<tsn::CsrItem> tsn::csr_items(const std::vector<uint32_t>& ptr, uint32_t* slots, uint32_t* counters) {
  std::vector<CsrItem> items;
  *slots = 0;
  *counters = 0;
  const uint32_t rows = static_cast<uint32_t>(ptr.size()) - 1;
  uint32_t t = 0;
  while (t < rows) {
    const uint32_t len = ptr[t + 1] - ptr[t];
    if (len > static_cast<uint32_t>(kCsrShortEntries)) {
      const uint32_t n = div_up(len, kCsrChunkEntries);
      const uint32_t slot0 = n > 1 ? *slots : 0, counter = n > 1 ? *counters : 0;
      if (n > 1) {
        *slots += n;
        *counters += 1;
      }
      for (uint32_t c = 0; c < n; ++c)
        items.push_back(CsrItem{t, t + 1, ptr[t] + static_cast<uint32_t>(static_cast<uint64_t>(len) * c / n),
                                ptr[t] + static_cast<uint32_t>(static_cast<uint64_t>(len) * (c + 1) / n), c, n,
                                slot0, counter});
      ++t;
      continue;
    }
    const uint32_t start = t;
    if (len == 0) {  // a run of empty rows
      while (t < rows && t - start < static_cast<uint32_t>(kCsrEmptyRunRows) && ptr[t + 1] == ptr[t]) ++t;
    } else {  // empty rows after the first stay in the run (cheap there)
      while (t < rows && t - start < static_cast<uint32_t>(kCsrRowsPerItem) &&
             ptr[t + 1] - ptr[t] <= static_cast<uint32_t>(kCsrShortEntries))
        ++t;
    }
    items.push_back(CsrItem{start, t, ptr[start], ptr[t], 0, 0, 0, 0});
  }
  return items;
}

std::vector<tsn::ChunkItem> tsn::long_chunks(const std::vector<uint32_t>& off, uint32_t problem, uint32_t slot_base,
                                             uint32_t counter_base, uint32_t* slots, uint32_t* counters) {
  std::vector<ChunkItem> out;
  uint32_t slot = slot_base, counter = counter_base;
  for (size_t g = 0; g + 1 < off.size(); ++g) {
    const uint32_t b = off[g], e = off[g + 1];
    if (e - b <= static_cast<uint32_t>(kSplitAboveTiles)) continue;
    const uint32_t n = (e - b + kSplitAboveTiles - 1) / kSplitAboveTiles;
    for (uint32_t c = 0; c < n; ++c)
      out.push_back({static_cast<int32_t>(g), static_cast<int32_t>(chunk_begin(b, e, n, c)),
                     static_cast<int32_t>(chunk_begin(b, e, n, c + 1)), c, n, slot, counter, problem});
    slot += n;
    ++counter;
  }
  *slots = slot - slot_base;
  *counters = counter - counter_base;
  return out;
}

namespace tsn {

// Dense-control tiles from the sparse operands, one thread per (tile, lane):
// lane r + 16q holds compressed row r, slots 4q .. 4q+3, i.e. the two kept
// values of column chunks 2q and 2q+1; the index byte holds each chunk's
// 4-bit code (positions lo = code & 3, hi = code >> 2 of its two values in
// the chunk).  Out: expanded row r, columns 8q .. 8q+7 as {re 8 x f16},
// {im 8 x f16}, zeros elsewhere (the kept value written second wins if a
// code repeats a position, as in the reference's expansion).
__global__ __launch_bounds__(256) void expand_dense_kernel(const uint4* __restrict__ a,
                                                           const uint16_t* __restrict__ idx, uint64_t lanes,
                                                           uint4* __restrict__ dense) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= lanes) return;
  const uint4 v = a[e];
  const uint32_t code = idx[e];
  const uint32_t vals[2][4] = {{v.x & 0xFFFFu, v.x >> 16, v.y & 0xFFFFu, v.y >> 16},
                               {v.z & 0xFFFFu, v.z >> 16, v.w & 0xFFFFu, v.w >> 16}};
  uint4 out[2];
#pragma unroll
  for (int part = 0; part < 2; ++part) {  // re, im
    uint16_t d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t byte = (code >> (8 * part)) & 0xFFu;
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const uint32_t nib = (byte >> (4 * cc)) & 0xFu;
      d[4 * cc + (nib & 3u)] = static_cast<uint16_t>(vals[part][2 * cc]);
      d[4 * cc + (nib >> 2)] = static_cast<uint16_t>(vals[part][2 * cc + 1]);
    }
    out[part] = make_uint4(d[0] | (static_cast<uint32_t>(d[1]) << 16), d[2] | (static_cast<uint32_t>(d[3]) << 16),
                           d[4] | (static_cast<uint32_t>(d[5]) << 16), d[6] | (static_cast<uint32_t>(d[7]) << 16));
  }
  dense[2 * e] = out[0];
  dense[2 * e + 1] = out[1];
}

int ensure_dense_tiles(const tsn_direction* cd) {
  if (!cd || cd->csr) return fail(TSN_ERR_ARGUMENT, "dense tiles need a packed (non-CSR) direction");
  if (cd->tile_dense) return TSN_OK;
  if (cd->h.tiles > kMaxDenseTiles)
    return fail(TSN_ERR_ARGUMENT, "pack has %u tiles, dense-control tiles hold at most %u", cd->h.tiles,
                kMaxDenseTiles);
  // the dense tiles are a cache of the direction's own operands
  auto* d = const_cast<tsn_direction*>(cd);
  const uint64_t lanes = static_cast<uint64_t>(d->h.tiles) * 64;
  if (lanes == 0) return TSN_OK;
  uint4* dense = nullptr;
  TSN_HIP(hipMalloc(reinterpret_cast<void**>(&dense), lanes * 2 * sizeof(uint4)));
  hipLaunchKernelGGL(expand_dense_kernel, dim3(div_up(lanes, 256)), dim3(256), 0, 0, d->tile_a, d->tile_idx, lanes,
                     dense);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(dense);
    return fail(TSN_ERR_RUNTIME, "dense tile expansion: %s", hipGetErrorString(e));
  }
  d->tile_dense = dense;
  d->bytes += lanes * 2 * sizeof(uint4);
  return TSN_OK;
}

// Chunk table, partial slots and zeroed arrival counters of a direction's
// long groups (fixed by the pack; rescheduling keeps them).
static int upload_chunks(tsn_direction* d) {
  uint32_t slots = 0, counters = 0;
  const std::vector<ChunkItem> chunks = long_chunks(d->host_offsets, 0, 0, 0, &slots, &counters);
  if (chunks.empty()) return TSN_OK;
  if (slots > kMaxSlots) return fail(TSN_ERR_ARGUMENT, "%u long-group chunks (at most %u)", slots, kMaxSlots);
  size_t bytes = 0;
  TSN_CHECK(device_copy(chunks.data(), chunks.size(), &d->chunks, &bytes));
  d->n_chunks = static_cast<uint32_t>(chunks.size());
  d->n_slots = slots;
  d->n_counters = counters;
  TSN_HIP(hipMalloc(reinterpret_cast<void**>(&d->partials), static_cast<size_t>(slots) * kSlotFloats * 4));
  TSN_HIP(hipMalloc(reinterpret_cast<void**>(&d->counters), counters * sizeof(uint32_t)));
  TSN_HIP(hipMemset(d->counters, 0, counters * sizeof(uint32_t)));
  d->bytes += bytes + static_cast<size_t>(slots) * kSlotFloats * 4 + counters * sizeof(uint32_t);
  return TSN_OK;
}

static int upload_pack(const HostPack& p, tsn_direction** out) {
  const tsn_pack_header& h = p.h;
  if (h.tiles > kMaxDeviceTiles)
    return fail(TSN_ERR_ARGUMENT, "pack has %u tiles, the device format holds at most %u", h.tiles, kMaxDeviceTiles);
  TSN_CHECK(validate_pack(p));
  const size_t tiles = h.tiles;
  std::vector<uint4> tile_a(tiles * 64);
  std::vector<uint16_t> tile_idx(tiles * 64);
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
    }
  }
  // Device column words (see TileCols): byte offset of the column's row in a
  // half panel, bit 31 of word 2 of each lane segment = last tile of a group.
  if (h.n_cols > kMaxDeviceCols)
    return fail(TSN_ERR_ARGUMENT, "pack has %u columns, the device format holds at most %u", h.n_cols,
                kMaxDeviceCols);
  std::vector<uint32_t> dev_cols(p.tile_col_ids.size());
  for (size_t i = 0; i < dev_cols.size(); ++i) dev_cols[i] = static_cast<uint32_t>(p.tile_col_ids[i]) * 32u;
  for (uint32_t g = 0; g < h.groups; ++g) {
    if (p.group_offsets[g + 1] == p.group_offsets[g]) continue;
    const size_t t = p.group_offsets[g + 1] - 1;
    for (int q = 0; q < 4; ++q) dev_cols[t * kK + q * 8 + 2] |= 0x80000000u;
  }
  auto* d = new tsn_direction();
  d->h = h;
  d->values = p.values;
  (void)hipGetDevice(&d->device);
  d->host_offsets = p.group_offsets;
  size_t bytes = 0;
  int rc = TSN_OK;
  if (rc == TSN_OK) rc = device_copy(p.group_offsets.data(), p.group_offsets.size(), &d->group_offsets, &bytes);
  if (rc == TSN_OK) rc = device_copy(p.group_row_ids.data(), p.group_row_ids.size(), &d->group_rows, &bytes);
  if (rc == TSN_OK) rc = device_copy(dev_cols.data(), dev_cols.size(), &d->tile_cols, &bytes);
  if (rc == TSN_OK) rc = device_copy(tile_a.data(), tile_a.size(), &d->tile_a, &bytes);
  if (rc == TSN_OK) rc = device_copy(tile_idx.data(), tile_idx.size(), &d->tile_idx, &bytes);
  d->bytes = bytes;
  if (rc == TSN_OK) rc = schedule_units(d, 0);
  if (rc == TSN_OK) rc = upload_chunks(d);
  if (rc != TSN_OK) {
    release(d);
    return rc;
  }
  *out = d;
  return TSN_OK;
}

constexpr uint32_t kMemberGroup = 8;  // default member group (batch autotuner candidate)

int interp_batch_build(const std::vector<InterpBatchPart>& parts, InterpBatch* out, uint32_t member_group,
                       uint32_t coil_waves, uint32_t depth) {
  *out = InterpBatch{};
  if (depth > 3) return fail(TSN_ERR_ARGUMENT, "pipeline depth must be 0 (default) or 1..3");
  if (member_group == 0) member_group = kMemberGroup;
  if (coil_waves != 1 && coil_waves != 2) return fail(TSN_ERR_ARGUMENT, "coil_waves must be 1 or 2");
  if (parts.empty()) return fail(TSN_ERR_ARGUMENT, "empty interpolation batch");
  const int values = parts[0].dir->values;
  const uint32_t blocks = parts[0].blocks;
  if (blocks != 1 && blocks != 2) return fail(TSN_ERR_ARGUMENT, "a batch part covers 1 or 2 coil blocks");
  std::vector<InterpProblem> probs;
  for (const InterpBatchPart& p : parts) {
    const tsn_direction* d = p.dir;
    if (!d || d->csr) return fail(TSN_ERR_ARGUMENT, "batched interpolation needs packed directions");
    if (d->values != values) return fail(TSN_ERR_ARGUMENT, "a batch mixes tile value formats");
    if (p.blocks != blocks) return fail(TSN_ERR_ARGUMENT, "a batch mixes parts of 1 and 2 coil blocks");
    InterpProblem q{};
    q.group_rows = d->group_rows;
    q.tile_cols = d->tile_cols;
    q.tile_a = d->tile_a;
    q.tile_idx = d->tile_idx;
    q.tile_dense = d->tile_dense;
    q.n_tiles = d->h.tiles;
    q.n_groups = d->h.groups;
    q.n_out = d->h.output_size;
    q.n_cols = d->h.n_cols;
    q.coils = p.coils;
    q.coil_base = p.coil_base;
    q.panel_il = p.panel_il;
    q.out_f32 = p.out_f32;
    q.out_il = p.out_il;
    // row-major output: compact rows of the problem's coils (one block), or
    // 64-byte rows in blocks of 8 coils (two blocks)
    q.out_rs = blocks == 2 ? kCoilsPerBlock : p.coils;
    q.out_bs = blocks == 2 ? static_cast<uint64_t>(q.n_out) * kCoilsPerBlock : kCoilsPerBlock;
    probs.push_back(q);
  }
  // Parts are grouped by direction, in order of first appearance.  Each
  // direction's own launch order (cut into one range per XCD, longest first)
  // is split at its XCD boundaries, and XCD x of the batch runs piece x of
  // every direction, one after the other: as in a launch of its own, each
  // XCD's L2 holds one eighth of a panel's columns at a time.  Within a
  // piece, groups of kMemberGroup members of a direction (frames sharing a
  // pack) follow each entry, so they reuse its tile operands from L2 while
  // only that many panels are gathered from at once (24 frames of 768^2 / 16
  // coils: 243 ms with all members per entry, 231 ms in groups of 8).
  std::vector<const tsn_direction*> dirs;
  for (const InterpBatchPart& p : parts)
    if (std::find(dirs.begin(), dirs.end(), p.dir) == dirs.end()) dirs.push_back(p.dir);
  std::vector<WorkItem> items;
  std::vector<uint32_t> item_problem;
  std::vector<ChunkItem> chunks;
  std::vector<uint32_t> xcd_entries[8];
  uint32_t slots = 0, counters = 0;
  for (const tsn_direction* d : dirs) {
    std::vector<uint32_t> members;
    for (uint32_t i = 0; i < parts.size(); ++i)
      if (parts[i].dir == d) members.push_back(i);
    // every part has its own slots and counters
    uint32_t ns = 0, nc = 0;
    const std::vector<ChunkItem> dc = long_chunks(d->host_offsets, 0, 0, 0, &ns, &nc);
    const uint32_t nd = static_cast<uint32_t>(d->host_order.size());
    for (uint32_t x = 0; x < 8; ++x) {
      const uint32_t lo = xcd_remap(x, nd), hi = x == 7 ? nd : xcd_remap(x + 1, nd);
      const uint32_t nm = static_cast<uint32_t>(members.size());
      for (uint32_t k0 = 0; k0 < nm; k0 += member_group)
      for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t e = d->host_order[i];
        for (uint32_t k = k0; k < std::min(nm, k0 + member_group); ++k) {
          if (e & kChunkTag) {
            ChunkItem c = dc[e & ~kChunkTag];
            c.slot0 += slots + k * ns;
            c.counter += counters + k * nc;
            c.problem = members[k];
            xcd_entries[x].push_back(kChunkTag | static_cast<uint32_t>(chunks.size()));
            chunks.push_back(c);
          } else {
            xcd_entries[x].push_back(static_cast<uint32_t>(items.size()));
            items.push_back(d->host_items[e]);
            item_problem.push_back(members[k]);
          }
        }
      }
    }
    slots += static_cast<uint32_t>(members.size()) * ns;
    counters += static_cast<uint32_t>(members.size()) * nc;
  }
  // The launch takes XCD ranges of equal length (xcd_remap), so a few entries
  // at each boundary run on the neighbouring XCD.
  std::vector<uint32_t> order;
  for (const std::vector<uint32_t>& v : xcd_entries) order.insert(order.end(), v.begin(), v.end());
  InterpBatch b;
  b.blocks = blocks;
  b.coil_waves = coil_waves;
  b.depth = depth;
  b.member_group = member_group;
  b.values = values;
  b.dense = std::all_of(parts.begin(), parts.end(), [](const InterpBatchPart& p) { return p.dir->tile_dense; });
  b.n_problems = static_cast<uint32_t>(probs.size());
  b.n_items = static_cast<uint32_t>(items.size());
  b.n_chunks = static_cast<uint32_t>(chunks.size());
  size_t bytes = 0;
  InterpProblem* dprobs = nullptr;
  int rc = device_copy(probs.data(), probs.size(), &dprobs, &bytes);
  b.problems = dprobs;
  if (rc == TSN_OK) rc = device_copy(items.data(), items.size(), &b.items, &bytes);
  if (rc == TSN_OK) rc = device_copy(item_problem.data(), item_problem.size(), &b.item_problem, &bytes);
  if (rc == TSN_OK) rc = device_copy(chunks.data(), chunks.size(), &b.chunks, &bytes);
  if (rc == TSN_OK) rc = device_copy(order.data(), order.size(), &b.order, &bytes);
  if (rc == TSN_OK && slots > kMaxSlots)
    rc = fail(TSN_ERR_ARGUMENT, "%u long-group chunks in the batch (at most %u)", slots, kMaxSlots);
  if (rc == TSN_OK && slots) {
    const hipError_t e1 = hipMalloc(reinterpret_cast<void**>(&b.partials), static_cast<size_t>(slots) * kSlotFloats * 4);
    const hipError_t e2 = e1 == hipSuccess ? hipMalloc(reinterpret_cast<void**>(&b.counters), counters * 4) : e1;
    const hipError_t e3 = e2 == hipSuccess ? hipMemset(b.counters, 0, counters * 4) : e2;
    if (e3 != hipSuccess) rc = fail(TSN_ERR_RUNTIME, "batch chunk buffers: %s", hipGetErrorString(e3));
  }
  if (rc != TSN_OK) {
    interp_batch_free(&b);
    return rc;
  }
  *out = b;
  return TSN_OK;
}

void interp_batch_free(InterpBatch* b) {
  if (!b) return;
  void* ptrs[] = {b->problems, b->items, b->item_problem, b->chunks, b->partials, b->counters, b->order};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  *b = InterpBatch{};
}

template <int KIND, int NB, int WB = 1>
static void launch_batch_nb(const InterpArgs& a, int out_mode, bool conj, hipStream_t s) {
  if (out_mode == kOutRowMajor) launch_tiles_om<KIND, NB, true, kOutRowMajor, WB>(a, conj, false, s);
  else if (KIND == kSparseSplitBF16) launch_tiles_om<KIND, NB, true, kOutSplitBF16IL, WB>(a, conj, false, s);
  else launch_tiles_om<KIND, NB, true, kOutSplit16IL, WB>(a, conj, false, s);
}

template <int KIND>
static void launch_batch_kind(const InterpArgs& a, uint32_t blocks, uint32_t coil_waves, int out_mode, bool conj,
                              hipStream_t s) {
  if (blocks == 2 && coil_waves == 2) launch_batch_nb<KIND, 1, 2>(a, out_mode, conj, s);
  else if (blocks == 2) launch_batch_nb<KIND, 2>(a, out_mode, conj, s);
  else launch_batch_nb<KIND, 1>(a, out_mode, conj, s);
}

int launch_interp_batch(const InterpBatch& b, int kind, int out_mode, bool conj, float scale, hipStream_t s) {
  if (!b.problems) return fail(TSN_ERR_ARGUMENT, "empty interpolation batch");
  if (out_mode != kOutRowMajor && out_mode != kOutSplit16IL)
    return fail(TSN_ERR_ARGUMENT, "batched interpolation writes row-major float2 or split panels");
  if ((kind == kSparseSplitBF16) != (b.values == TSN_VALUES_BF16))
    return fail(TSN_ERR_ARGUMENT, "tile value format of the batch does not match the interpolation variant");
  if (kind == kDenseSplit && !b.dense) return fail(TSN_ERR_ARGUMENT, "batch has no dense-control tiles");
  InterpArgs a{};
  a.items = b.items;
  a.n_items = b.n_items;
  a.chunks = b.chunks;
  a.n_chunks = b.n_chunks;
  a.order = b.order;
  a.partials = b.partials;
  a.counters = b.counters;
  a.scale = scale;
  a.problems = static_cast<const InterpProblem*>(b.problems);
  a.item_problem = b.item_problem;
  a.coils = kCoilsPerBlock;  // replaced per problem
  a.depth = b.depth;
  switch (kind) {
    case kSparseSplit: launch_batch_kind<kSparseSplit>(a, b.blocks, b.coil_waves, out_mode, conj, s); break;
    case kDenseSplit: launch_batch_kind<kDenseSplit>(a, b.blocks, b.coil_waves, out_mode, conj, s); break;
    case kSparseHi: launch_batch_kind<kSparseHi>(a, b.blocks, b.coil_waves, out_mode, conj, s); break;
    case kSparseSplitBF16:
      launch_batch_kind<kSparseSplitBF16>(a, b.blocks, b.coil_waves, out_mode, conj, s);
      break;
    default: return fail(TSN_ERR_ARGUMENT, "unknown interpolation kind %d", kind);
  }
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

// ------------------------------------------------- no-allocation entry point
// tsn_interp_noalloc takes the long-group chunk scratch from this file-scope
// array instead of the direction's: kStaticSlots partial slots (4 KiB each)
// followed by kStaticCounters arrival counters, 6 MiB + 16 KiB in all.  Every
// device the library runs on holds its own copy until the process ends.  On
// the host `scratch` names a host-side stand-in, never the device memory: the
// device address comes from hipGetSymbolAddress, once per device, and only
// that address is handed to kernels.
constexpr uint32_t kStaticSlots = 1536;
constexpr uint32_t kStaticCounters = 4096;
constexpr size_t kStaticScratchWords = static_cast<size_t>(kStaticSlots) * kSlotFloats + kStaticCounters;
static __device__ __align__(256) uint32_t scratch[kStaticScratchWords];
constexpr int kMaxStaticDevices = 64;

// Device address of `scratch` on the current device (looked up on the first
// call per device).
static int static_scratch(uint32_t** out) {
  static std::mutex lock;
  static uint32_t* address[kMaxStaticDevices] = {};
  int device = 0;
  TSN_HIP(hipGetDevice(&device));
  if (device < 0 || device >= kMaxStaticDevices)
    return fail(TSN_ERR_ARGUMENT, "device %d is beyond the %d devices of the static scratch table", device,
                kMaxStaticDevices);
  std::lock_guard<std::mutex> guard(lock);
  if (!address[device]) {
    void* ptr = nullptr;
    TSN_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(TSN_ERR_RUNTIME, "no device address for the static scratch on device %d", device);
    address[device] = static_cast<uint32_t*>(ptr);
  }
  *out = address[device];
  return TSN_OK;
}

__global__ __launch_bounds__(256) void clear_words_kernel(uint4* __restrict__ p, uint32_t n4) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n4) p[e] = make_uint4(0u, 0u, 0u, 0u);
}

}  // namespace tsn

extern "C" {

int tsn_interp_noalloc(const tsn_direction* d, const void* hi, const void* res, void* out, uint32_t coils,
                       int conjugate, int accumulate, float scale, int out_layout, void* stream) {
  if (out_layout != TSN_OUT_COIL_MAJOR && out_layout != TSN_OUT_ROW_MAJOR)
    return fail(TSN_ERR_ARGUMENT, "unknown output layout %d", out_layout);
  if (!d || d->csr) return fail(TSN_ERR_ARGUMENT, "interpolation needs a packed (non-CSR) direction");
  if (d->n_slots > kStaticSlots || d->n_counters > kStaticCounters)
    return fail(TSN_ERR_ARGUMENT,
                "direction needs %u chunk slots and %u counters; the static scratch holds %u and %u",
                d->n_slots, d->n_counters, kStaticSlots, kStaticCounters);
  int device = 0;
  TSN_HIP(hipGetDevice(&device));
  if (device != d->device)
    return fail(TSN_ERR_ARGUMENT, "direction lives on device %d, the current device is %d", d->device, device);
  ChunkScratch chunk;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (d->n_chunks) {
    uint32_t* base = nullptr;
    TSN_CHECK(static_scratch(&base));
    chunk.partials = reinterpret_cast<float*>(base);
    chunk.counters = base + static_cast<size_t>(kStaticSlots) * kSlotFloats;
    // the counters this launch uses start at zero whatever the last call left
    const uint32_t n4 = div_up(d->n_counters, 4);
    hipLaunchKernelGGL(clear_words_kernel, dim3(div_up(n4, 256)), dim3(256), 0, s,
                       reinterpret_cast<uint4*>(chunk.counters), n4);
    TSN_HIP(hipGetLastError());
  }
  return launch_interp(d, res ? kSparseSplit : kSparseHi, hi, res, coils, out_layout, conjugate != 0,
                       accumulate != 0, scale, static_cast<float*>(out), nullptr, nullptr, s, false, nullptr,
                       d->n_chunks ? &chunk : nullptr);
}

int tsn_direction_upload(const tsn_host_pack* pack, tsn_direction** out) {
  if (!pack || !out) return fail(TSN_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  return upload_pack(*pack, out);
}

int tsn_direction_schedule(tsn_direction* d, uint32_t unit_tiles) {
  if (!d) return fail(TSN_ERR_ARGUMENT, "null direction");
  if (d->csr) return fail(TSN_ERR_ARGUMENT, "a CSR direction has no tile schedule");
  if (unit_tiles > (1u << 20)) return fail(TSN_ERR_ARGUMENT, "unit_tiles %u is out of range", unit_tiles);
  TSN_HIP(hipSetDevice(d->device));
  TSN_HIP(hipDeviceSynchronize());
  return schedule_units(d, unit_tiles);
}

uint32_t tsn_direction_unit_tiles(const tsn_direction* d) { return d ? d->unit_tiles : 0; }

int tsn_direction_set_coil_waves(tsn_direction* d, uint32_t waves) {
  if (!d) return fail(TSN_ERR_ARGUMENT, "null direction");
  if (waves != 1 && waves != 2) return fail(TSN_ERR_ARGUMENT, "coil waves must be 1 or 2 (got %u)", waves);
  d->coil_waves = waves;
  return TSN_OK;
}

uint32_t tsn_direction_coil_waves(const tsn_direction* d) { return d ? d->coil_waves : 0; }

int tsn_direction_set_pipe_depth(tsn_direction* d, uint32_t depth) {
  if (!d) return fail(TSN_ERR_ARGUMENT, "null direction");
  if (depth > 3) return fail(TSN_ERR_ARGUMENT, "pipeline depth must be 0 (default) or 1..3");
  d->pipe_depth = depth;
  return TSN_OK;
}

uint32_t tsn_direction_pipe_depth(const tsn_direction* d) { return d ? d->pipe_depth : 0; }

int tsn_direction_autotune(tsn_direction* d, uint32_t coils, void* stream, uint32_t* best_unit_tiles,
                           float* best_us) {
  if (!d) return fail(TSN_ERR_ARGUMENT, "null direction");
  if (d->csr) return fail(TSN_ERR_ARGUMENT, "a CSR direction has no tile schedule");
  if (coils == 0 || coils > static_cast<uint32_t>(kMaxCoils)) return fail(TSN_ERR_ARGUMENT, "coils must be in 1..16");
  hipStream_t s = static_cast<hipStream_t>(stream);
  TSN_HIP(hipSetDevice(d->device));
  const uint32_t blocks = div_up(coils, kCoilsPerBlock);
  const size_t panel_words = static_cast<size_t>(blocks) * d->h.n_cols * kN;
  // timed with the interleaved split output (64-byte rows per coil block, as
  // plans write both directions' outputs: split panels forward, float2 rows
  // adjoint); a coil-major output's scattered writes rank the units differently
  const size_t out_words = static_cast<size_t>(blocks) * d->h.output_size * kN;
  uint32_t* panel = nullptr;
  uint32_t* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = TSN_OK;
  auto hip = [&](hipError_t e, const char* what) {
    if (rc == TSN_OK && e != hipSuccess) rc = fail(TSN_ERR_RUNTIME, "autotune %s: %s", what, hipGetErrorString(e));
  };
  hip(hipMalloc(reinterpret_cast<void**>(&panel), panel_words * 4), "alloc");
  hip(hipMalloc(reinterpret_cast<void**>(&out), out_words * 4), "alloc");
  if (rc == TSN_OK) hip(hipMemsetAsync(panel, 0, panel_words * 4, s), "clear");
  hip(hipEventCreate(&e0), "event");
  hip(hipEventCreate(&e1), "event");
  const int kind = d->values == TSN_VALUES_BF16 ? kSparseSplitBF16 : kSparseSplit;
  // candidates: unit sizes x, for 9..16 coils, one wave for both coil blocks
  // or one per block x pipeline depths (the default and the alternatives of
  // the launch: 1 or 3 tiles ahead for one block per wave, 2 for two)
  struct Candidate {
    uint32_t waves, unit, depth;
  };
  std::vector<Candidate> cands;
  const bool two = coils > static_cast<uint32_t>(kCoilsPerBlock);
  for (uint32_t cw : two ? std::vector<uint32_t>{1u, 2u} : std::vector<uint32_t>{d->coil_waves})
    for (uint32_t u : {2u, 3u, 4u, 5u, 6u, 8u, 10u, 12u, 14u, 16u}) {
      std::vector<uint32_t> depths = {0u};
      if (two && cw == 1) depths.push_back(2u);
      if (!two) depths.insert(depths.end(), {1u, 3u});
      for (uint32_t dp : depths) cands.push_back({cw, u, dp});
    }
  // two passes over the candidates, each candidate's faster pass counts
  // (neighbouring unit sizes differ by a few percent, about the spread of
  // one timing)
  std::vector<float> times(cands.size(), 0.0f);
  Candidate best{d->coil_waves, 0, d->pipe_depth};
  float best_ms = 0.0f;
  for (int pass = 0; pass < 2; ++pass)
  for (size_t ci = 0; ci < cands.size(); ++ci) {
    if (rc != TSN_OK) break;
    const Candidate& c = cands[ci];
    d->coil_waves = c.waves;
    d->pipe_depth = c.depth;
    rc = schedule_units(d, c.unit);
    auto launch = [&]() {
      return launch_interp(d, kind, panel, nullptr, coils, kOutSplit16IL, false, false, 1.0f, nullptr, nullptr,
                           nullptr, s, true, out);
    };
    for (int i = 0; rc == TSN_OK && i < 3; ++i) rc = launch();
    if (rc != TSN_OK) break;
    hip(hipEventRecord(e0, s), "event");
    constexpr int kReps = 20;
    for (int i = 0; rc == TSN_OK && i < kReps; ++i) rc = launch();
    hip(hipEventRecord(e1, s), "event");
    hip(hipEventSynchronize(e1), "event");
    float ms = 0.0f;
    hip(hipEventElapsedTime(&ms, e0, e1), "event");
    ms /= kReps;
    if (rc != TSN_OK) break;
    if (pass == 1) ms = std::min(ms, times[ci]);
    times[ci] = ms;
    if (pass == 1 && (best.unit == 0 || ms < best_ms)) {
      best = c;
      best_ms = ms;
    }
  }
  (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (panel) (void)hipFree(panel);
  if (out) (void)hipFree(out);
  d->coil_waves = best.waves;
  d->pipe_depth = best.depth;
  if (rc == TSN_OK) rc = schedule_units(d, best.unit);
  if (rc != TSN_OK) return rc;
  if (best_unit_tiles) *best_unit_tiles = best.unit;
  if (best_us) *best_us = best_ms * 1000.0f;
  return TSN_OK;
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
  // Rows in Morton order of their k-space position (samples: the rounded grid
  // coordinate; grid points: their own), so the contiguous range of work
  // items each XCD runs covers a compact k-space region and its gathers stay
  // in that XCD's L2.  Entries keep their order, so the sums do not change.
  {
    const uint32_t n = static_cast<uint32_t>(rows.size());
    std::vector<uint32_t> key(n);
    for (uint32_t t = 0; t < n; ++t) {
      if (view == 0) {
        const double c0 = std::nearbyint(kb->coord0[t]), c1 = std::nearbyint(kb->coord1[t]);
        key[t] = morton_key(static_cast<uint32_t>(std::max(0.0, c0)), static_cast<uint32_t>(std::max(0.0, c1)));
      } else {
        const uint32_t g = static_cast<uint32_t>(rows[t]);
        key[t] = morton_key(g % kb->grid1, g / kb->grid1);
      }
    }
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    std::vector<uint32_t> nptr(1, 0);
    std::vector<int32_t> nrows(n), ncols(cols.size());
    std::vector<float2> nvals(vals.size());
    nptr.reserve(n + 1);
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t t = perm[i];
      nrows[i] = rows[t];
      const uint32_t b = ptr[t], e = ptr[t + 1], o = nptr.back();
      std::copy(cols.begin() + b, cols.begin() + e, ncols.begin() + o);
      std::copy(vals.begin() + b, vals.begin() + e, nvals.begin() + o);
      nptr.push_back(o + (e - b));
    }
    ptr.swap(nptr);
    rows.swap(nrows);
    cols.swap(ncols);
    vals.swap(nvals);
  }
  // Output rows no entry reaches (grid points of the transpose no sample
  // touches) follow as empty CSR rows in ascending order, so a launch that
  // does not accumulate writes zeros there too (runs of empty rows).  Their
  // items come last; a launch into an output already zero leaves them out.
  const uint32_t active = static_cast<uint32_t>(rows.size());
  if (active < h.output_size) {
    std::vector<uint8_t> hit(h.output_size, 0);
    for (int32_t r : rows) hit[r] = 1;
    for (uint32_t r = 0; r < h.output_size; ++r)
      if (!hit[r]) {
        rows.push_back(static_cast<int32_t>(r));
        ptr.push_back(ptr.back());
      }
  }
  uint32_t slots = 0, counters = 0;
  const std::vector<CsrItem> items = csr_items(ptr, &slots, &counters);
  uint32_t active_items = 0;  // items holding an active row: all but the trailing runs of empty rows
  while (active_items < items.size() && items[active_items].row_begin < active) ++active_items;
  if (slots > kMaxCsrSlots) return fail(TSN_ERR_ARGUMENT, "%u long-row chunks (at most %u)", slots, kMaxCsrSlots);
  auto* d = new tsn_direction();
  d->h = h;
  d->csr = true;
  (void)hipGetDevice(&d->device);
  d->csr_rows = static_cast<uint32_t>(rows.size());
  d->n_csr_items = static_cast<uint32_t>(items.size());
  d->n_csr_active_items = active_items;
  d->n_slots = slots;
  d->n_counters = counters;
  size_t bytes = 0;
  int rc = device_copy(ptr.data(), ptr.size(), &d->row_ptr, &bytes);
  if (rc == TSN_OK) rc = device_copy(rows.data(), rows.size(), &d->row_ids, &bytes);
  if (rc == TSN_OK) rc = device_copy(cols.data(), cols.size(), &d->cols, &bytes);
  if (rc == TSN_OK) rc = device_copy(vals.data(), vals.size(), &d->vals, &bytes);
  if (rc == TSN_OK) rc = device_copy(items.data(), items.size(), &d->csr_items, &bytes);
  if (rc == TSN_OK && slots) {
    const std::vector<float> zeros(static_cast<size_t>(slots) * 2 * kMaxCoils, 0.0f);
    const std::vector<uint32_t> zero_counters(round_up(counters, 4u), 0u);
    rc = device_copy(zeros.data(), zeros.size(), &d->partials, &bytes);
    if (rc == TSN_OK) rc = device_copy(zero_counters.data(), zero_counters.size(), &d->counters, &bytes);
  }
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

int tsn_direction_expand_dense(tsn_direction* d) { return ensure_dense_tiles(d); }

void tsn_direction_free(tsn_direction* d) { release(d); }

int tsn_interp(const tsn_direction* d, const void* hi, const void* res, void* out, uint32_t coils,
               int conjugate, int accumulate, float scale, int out_layout, void* stream) {
  if (out_layout != TSN_OUT_COIL_MAJOR && out_layout != TSN_OUT_ROW_MAJOR)
    return fail(TSN_ERR_ARGUMENT, "unknown output layout %d", out_layout);
  return launch_interp(d, res ? kSparseSplit : kSparseHi, hi, res, coils, out_layout, conjugate != 0,
                       accumulate != 0, scale, static_cast<float*>(out), nullptr, nullptr,
                       static_cast<hipStream_t>(stream), false, nullptr);
}

int tsn_interp_dense(const tsn_direction* d, const void* hi, const void* res, void* out,
                     uint32_t coils, int conjugate, int accumulate, float scale, int out_layout,
                     void* stream) {
  if (out_layout != TSN_OUT_COIL_MAJOR && out_layout != TSN_OUT_ROW_MAJOR)
    return fail(TSN_ERR_ARGUMENT, "unknown output layout %d", out_layout);
  if (!res) return fail(TSN_ERR_ARGUMENT, "dense control needs hi and residual panels");
  TSN_CHECK(ensure_dense_tiles(d));
  return launch_interp(d, kDenseSplit, hi, res, coils, out_layout, conjugate != 0, accumulate != 0,
                       scale, static_cast<float*>(out), nullptr, nullptr,
                       static_cast<hipStream_t>(stream), false, nullptr);
}

int tsn_interp_f32(const tsn_direction* d, const void* in, void* out, uint32_t coils, int conjugate,
                   int accumulate, float scale, int out_layout, void* stream) {
  return launch_interp_csr(d, static_cast<const float2*>(in), coils, out_layout, conjugate != 0,
                           accumulate != 0, scale, static_cast<float2*>(out),
                           static_cast<hipStream_t>(stream));
}

int tsn_interp_il(const tsn_direction* d, const void* panel, void* out, uint32_t coils, int conjugate,
                  int accumulate, float scale, int out_layout, void* stream) {
  if (out_layout != TSN_OUT_COIL_MAJOR && out_layout != TSN_OUT_ROW_MAJOR && out_layout != TSN_OUT_SPLIT_IL)
    return fail(TSN_ERR_ARGUMENT, "unknown output layout %d", out_layout);
  if (!d) return fail(TSN_ERR_ARGUMENT, "null direction");
  const int kind = d->values == TSN_VALUES_BF16 ? kSparseSplitBF16 : kSparseSplit;
  if (out_layout == TSN_OUT_SPLIT_IL) {
    if (accumulate) return fail(TSN_ERR_ARGUMENT, "a split-panel output cannot accumulate");
    return launch_interp(d, kind, panel, nullptr, coils, kOutSplit16IL, conjugate != 0, false, scale, nullptr,
                         nullptr, nullptr, static_cast<hipStream_t>(stream), true, static_cast<uint32_t*>(out));
  }
  return launch_interp(d, kind, panel, nullptr, coils, out_layout, conjugate != 0, accumulate != 0, scale,
                       static_cast<float*>(out), nullptr, nullptr, static_cast<hipStream_t>(stream), true, nullptr);
}

int tsn_interp_dc(const tsn_direction* d, const void* panel, const void* measurements, int measurement_layout,
                  const float* density, uint32_t coils, float scale, float measurement_scale, int mode,
                  void* workspace, void* out_panel, void* stream) {
  if (!d || !panel || !measurements || !out_panel) return fail(TSN_ERR_ARGUMENT, "null argument");
  if (d->csr) return fail(TSN_ERR_ARGUMENT, "data consistency needs a packed (non-CSR) direction");
  if (coils == 0 || coils > static_cast<uint32_t>(kMaxCoils)) return fail(TSN_ERR_ARGUMENT, "coils must be in 1..16");
  if (measurement_layout != TSN_OUT_COIL_MAJOR && measurement_layout != TSN_OUT_ROW_MAJOR)
    return fail(TSN_ERR_ARGUMENT, "unknown measurement layout %d", measurement_layout);
  if (mode != TSN_DC_MATERIALIZED && mode != TSN_DC_FUSED)
    return fail(TSN_ERR_ARGUMENT, "data-consistency mode must be 1 (materialized) or 2 (fused), got %d", mode);
  if (mode == TSN_DC_MATERIALIZED && !workspace) return fail(TSN_ERR_ARGUMENT, "materialized mode needs a workspace");
  const uint32_t rows = d->h.output_size;
  DcTerms dc;
  dc.y = static_cast<const float*>(measurements);
  dc.density = density;
  dc.scale = measurement_scale;
  dc.row_stride = measurement_layout == TSN_OUT_ROW_MAJOR ? 2 * coils : 2;
  dc.coil_stride = measurement_layout == TSN_OUT_ROW_MAJOR ? 2 : 2 * rows;
  const bool bf16 = d->values == TSN_VALUES_BF16;
  const int kind = bf16 ? kSparseSplitBF16 : kSparseSplit;
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto* out = static_cast<uint32_t*>(out_panel);
  if (mode == TSN_DC_FUSED) return launch_interp_dc(d, kind, static_cast<const uint32_t*>(panel), coils, scale, dc, out, s);
  TSN_CHECK(launch_interp(d, kind, panel, nullptr, coils, kOutRowMajor, false, false, scale,
                          static_cast<float*>(workspace), nullptr, nullptr, s, true, nullptr));
  return launch_dc_residual(static_cast<const float2*>(workspace), rows, coils, dc, bf16, out, s);
}

int tsn_split_panel_il(const void* in, uint32_t rows, uint32_t coils, float scale, int values, void* panel,
                       void* stream) {
  return launch_split_panel_il(static_cast<const float2*>(in), rows, coils, scale,
                               static_cast<uint32_t*>(panel), static_cast<hipStream_t>(stream), values);
}

int tsn_split_panel(const void* in, uint32_t rows, uint32_t coils, float scale, void* hi, void* res,
                    void* stream) {
  return launch_split_panel(static_cast<const float2*>(in), rows, coils, scale,
                            static_cast<_Float16*>(hi), static_cast<_Float16*>(res),
                            static_cast<hipStream_t>(stream));
}

}  // extern "C"
