// Internal declarations of the gfx950 NUFFT / CG-SENSE library (libtsn.so).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tsn.h"

namespace tsn {

constexpr uint64_t kMagic = 0x4E55464654323431ull;  // "NUFFT241"
constexpr uint32_t kVersion = 2;
constexpr int kM = 16;      // rows per group / tile
constexpr int kK = 32;      // columns per tile
constexpr int kKComp = 16;  // kept values per tile row
constexpr int kN = 16;      // channels per panel block (8 coils x Re/Im)
constexpr int kCoilsPerBlock = 8;
constexpr int kMaxCoils = 16;
constexpr int kSplitAboveTiles = 16;  // tiles per unit at most; longer groups are cut into chunks of this size
// Interleaved-panel interpolation of 9..16 coils: waves per work entry, 1 (one
// wave computes both coil blocks) or 2 (one wave per block, same entry).
constexpr uint32_t kDefaultCoilWaves = 1;
// FP32 CSR work items: runs of at most kCsrRowsPerItem consecutive rows of at
// most kCsrShortEntries entries each, the first not empty; longer rows are
// cut into chunks of at most kCsrChunkEntries entries; a run that would start
// with an empty row takes the consecutive empty rows instead, at most
// kCsrEmptyRunRows (no entries to stage).  A wave stages an item's entries in
// LDS.
constexpr int kCsrRowsPerItem = 8;
constexpr int kCsrEmptyRunRows = 64;
constexpr int kCsrShortEntries = 64;
constexpr int kCsrChunkEntries = 256;
constexpr int kCsrStageEntries = kCsrRowsPerItem * kCsrShortEntries;
static_assert(kCsrChunkEntries <= kCsrStageEntries, "a chunk must fit the LDS stage");
// Device column words of packed tiles hold column * 32 (a byte offset) below
// bit 31, so a pack may have at most 2^26 columns.
constexpr uint32_t kMaxDeviceCols = 1u << 26;
// Per-tile operand offsets are signed 32-bit buffer offsets: sparse operands
// (1 KiB per tile) hold 2^21 tiles, expanded dense-control tiles (2 KiB) 2^20.
constexpr uint32_t kMaxDeviceTiles = 1u << 21;
constexpr uint32_t kMaxDenseTiles = 1u << 20;

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define TSN_HIP(expr)                                                                   \
  do {                                                                                  \
    hipError_t tsn_status_ = (expr);                                                    \
    if (tsn_status_ != hipSuccess)                                                      \
      return ::tsn::fail(TSN_ERR_RUNTIME, "HIP error '%s' at %s:%d",                    \
                         hipGetErrorString(tsn_status_), __FILE__, __LINE__);           \
  } while (0)

#define TSN_CHECK(expr)                  \
  do {                                   \
    int tsn_rc_ = (expr);                \
    if (tsn_rc_ != TSN_OK) return tsn_rc_; \
  } while (0)

inline uint32_t div_up(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }
inline uint32_t round_up(uint32_t a, uint32_t b) { return (a + b - 1) / b * b; }

// Block b of a launch runs on XCD b % 8 (the dispatcher's round robin); this
// gives each XCD a contiguous range of the nb work indices (a permutation of
// 0 .. nb-1), so neighbouring work shares that XCD's L2.
__host__ __device__ inline uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t xcd = b & 7u, idx = b >> 3;
  const uint32_t per = nb >> 3, rem = nb & 7u;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// Morton (Z-order) key of a 2-D point with 16-bit coordinates.
inline uint32_t morton_key(uint32_t x, uint32_t y) {
  auto spread = [](uint32_t v) {
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
  };
  return spread(x) | (spread(y) << 1);
}

// bfloat16 bits of v, round to nearest even: the top half of the float32 bits
// after adding 0x7FFF plus the lowest kept bit; NaN stays a quiet NaN.
__host__ __device__ inline uint16_t bf16_rne_bits(float v) {
  uint32_t u = __builtin_bit_cast(uint32_t, v);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__host__ __device__ inline float bf16_to_float(uint16_t b) {
  return __builtin_bit_cast(float, static_cast<uint32_t>(b) << 16);
}

// One packed direction in host memory (real + imag share the structure).
struct HostPack {
  tsn_pack_header h{};
  std::vector<uint32_t> group_offsets;  // [groups + 1]
  std::vector<int32_t> group_row_ids;   // [groups * 16]
  std::vector<int32_t> tile_col_ids;    // [tiles * 32]
  int values = TSN_VALUES_FP16;         // format of the 16-bit tile values (files hold FP16)
  std::vector<uint16_t> real_comp;      // [tiles * 16 * 16] value bits
  std::vector<uint32_t> real_meta;      // [tiles * 16]
  std::vector<uint16_t> imag_comp;
  std::vector<uint32_t> imag_meta;
};

int validate_header(const tsn_pack_header& h);
int validate_pack(const HostPack& p);

// Decode the pair-packed metadata of one tile into 16 row words
// (reference experiments/packing/prepare_dense_control.py:27-36).
inline void decode_row_meta(const uint32_t* pair_meta, uint32_t* row_meta) {
  for (int g = 0; g < 8; ++g) {
    const uint32_t a = pair_meta[2 * g], b = pair_meta[2 * g + 1];
    row_meta[g] = (a & 0xFFFFu) | ((b & 0xFFFFu) << 16);
    row_meta[g + 8] = (a >> 16) | (b & 0xFFFF0000u);
  }
}

// Work unit of the interpolation kernels: whole groups [group, group_end)
// whose tiles [tile_begin, tile_end) one wave streams through a continuous
// pipeline (epilogue at every group boundary).  Units hold at most a target
// tile count; groups longer than kSplitAboveTiles are cut into chunks.
struct WorkItem {
  int32_t group;
  int32_t group_end;
  int32_t tile_begin;
  int32_t tile_end;
};

// Chunk c of the n = ceil(tiles / kSplitAboveTiles) chunks of a long group
// [b, e): tiles [chunk_begin(c), chunk_begin(c + 1)), sizes within one of each
// other, so no chunk is much longer than its group needs.
inline uint32_t chunk_begin(uint32_t b, uint32_t e, uint32_t n, uint32_t c) {
  return b + static_cast<uint32_t>((static_cast<uint64_t>(e - b) * c) / n);
}
// Launch-order rank of a chunk's finishing chain (publish, count, sum the
// slots), in tiles: chunks sort ahead of units of their own length.
constexpr int32_t kChunkFinishTiles = 4;

// One chunk of a long group: consecutive tiles of it (chunk_begin).  Each
// chunk's wave publishes its partial sums to slot
// slot0 + chunk; the wave that brings the group's arrival counter to
// n_chunks adds the slots in chunk order and writes the rows, so the sums do
// not depend on the schedule or on which wave arrives last.
struct ChunkItem {
  int32_t group;
  int32_t tile_begin, tile_end;
  uint32_t chunk, n_chunks;
  uint32_t slot0;    // partial slot of the group's chunk 0
  uint32_t counter;  // arrival counter of the group (reset to 0 by the last arrival)
  uint32_t problem;  // batched launches: the problem table entry
};
// One work item of the FP32 CSR kernel: a run of short rows (n_chunks 0), or
// chunk `chunk` of the n_chunks chunks of one long row.  A row of several
// chunks has partial slots slot0 .. slot0 + n_chunks - 1 (kMaxCoils float2
// each) and arrival counter `counter`.
struct CsrItem {
  uint32_t row_begin, row_end;  // CSR rows (a chunk: row_end = row_begin + 1)
  uint32_t e_begin, e_end;      // the item's entries, contiguous in the CSR arrays
  uint32_t chunk, n_chunks;
  uint32_t slot0, counter;
};
// Floats of one partial slot: two coil blocks x (re, im) x 64 lanes x float4.
constexpr uint32_t kSlotFloats = 2 * 2 * 64 * 4;
// Launch order entries: a unit index, or kChunkTag | chunk index.  Units and
// chunks are cut in pack order, so each XCD's contiguous range of the launch
// holds neighbouring groups (shared columns in its L2) and its share of the
// long groups; within windows of a range the longest entries go first.
constexpr uint32_t kChunkTag = 0x80000000u;
// Chunks of a direction's long groups with slots and counters numbered from
// slot_base / counter_base (host side).
// FP32 CSR work items of row pointers ptr (host side); slots and counters of
// the rows of several chunks are numbered from 0.
std::vector<CsrItem> csr_items(const std::vector<uint32_t>& ptr, uint32_t* slots, uint32_t* counters);
// Slot offsets of CSR chunks are 32-bit buffer offsets of kMaxCoils float2.
constexpr uint32_t kMaxCsrSlots = (1u << 31) / (kMaxCoils * 8);
std::vector<ChunkItem> long_chunks(const std::vector<uint32_t>& group_offsets, uint32_t problem, uint32_t slot_base,
                                   uint32_t counter_base, uint32_t* slots, uint32_t* counters);
// One part of a batched interpolation launch: one or two coil blocks of one
// problem (with two, one wave computes both from the same tile operands).
struct InterpBatchPart {
  const tsn_direction* dir;
  const uint32_t* panel_il;  // the first block's interleaved input panel [n_cols][16]
  float* out_f32;            // the problem's float2 output (row-major: compact rows of its
                             // coils up to 8 coils, 8-coil blocks of rows above 8)
  uint32_t* out_il;          // the first block's interleaved split output panel [n_out][16]
  uint32_t coils;            // coils of the problem
  uint32_t coil_base;        // first coil of the first block
  uint32_t blocks = 1;       // coil blocks of the part (1 or 2; block 1's panels follow block 0's)
};

// Device tables of a batched launch: XCD x runs piece x of every direction's
// own launch order, and groups of the parts sharing a direction follow each
// entry (so its tiles are read from HBM once per group, then from cache).
struct InterpBatch {
  void* problems = nullptr;
  WorkItem* items = nullptr;
  uint32_t* item_problem = nullptr;
  ChunkItem* chunks = nullptr;   // long-group chunks of every problem
  float* partials = nullptr;     // their partial slots
  uint32_t* counters = nullptr;  // one per problem x long group, zero between launches
  uint32_t* order = nullptr;     // launch order of units and chunks (kChunkTag marks a chunk)
  uint32_t n_problems = 0, n_items = 0, n_chunks = 0;
  uint32_t blocks = 1;  // coil blocks per part (every part of a batch has the same)
  uint32_t coil_waves = kDefaultCoilWaves;  // parts of two blocks: 1 wave for both, or 2 (one per block)
  uint32_t depth = 0;                       // tiles in flight ahead of the multiply (0: the kernel's default)
  uint32_t member_group = 0;                // parts sharing a direction that follow each entry
  int values = TSN_VALUES_FP16;
  bool dense = false;
};
// member_group: parts sharing a direction that follow each of its entries
// (0 = the default, 8); coil_waves: waves per entry of two-block parts;
// depth: tiles in flight ahead of the multiply (0 = the kernel's default).
int interp_batch_build(const std::vector<InterpBatchPart>& parts, InterpBatch* out, uint32_t member_group = 0,
                       uint32_t coil_waves = kDefaultCoilWaves, uint32_t depth = 0);
void interp_batch_free(InterpBatch* b);
// kind: 0 split FP16, 1 dense control, 2 FP16 hi only, 3 split BF16; out_mode:
// 1 row-major float2, 3 interleaved split (FP16 or BF16 as the kind).
int launch_interp_batch(const InterpBatch& b, int kind, int out_mode, bool conj, float scale, hipStream_t s);

// Scratch of the long-group chunk reduction: partial slots [n_slots][kSlotFloats]
// and arrival counters [n_counters] (zero between launches; the last arrival
// of a group resets its counter).  Launches that run at the same time need
// separate scratch: a direction holds one set for tsn_interp, and every plan
// holds its own for each of its directions.
struct ChunkScratch {
  float* partials = nullptr;
  uint32_t* counters = nullptr;
};
// FP32 CSR interpolation; scratch null: the direction's own chunk scratch.
// out_prezeroed: the output's empty rows hold zeros already (a plan's grid,
// zeroed at creation and written only here), so they are not written again.
int launch_interp_csr(const tsn_direction* d, const float2* in, uint32_t coils, int out_mode, bool conj,
                      bool accumulate, float scale, float2* out, hipStream_t stream,
                      const ChunkScratch* scratch = nullptr, bool out_prezeroed = false);
// Packed interpolation (kind and out_mode as below); scratch null: the
// direction's own.
int launch_interp(const tsn_direction* d, int kind, const void* hi, const void* res, uint32_t coils,
                  int out_mode, bool conj, bool accumulate, float scale, float* out_f32,
                  _Float16* out_hi, _Float16* out_res, hipStream_t stream, bool interleaved,
                  uint32_t* out_il, const ChunkScratch* scratch = nullptr);

// Data-consistency terms: residual (prediction - scale * y) * w per sample
// row, y read at row * row_stride + coil * coil_stride + part (floats).
struct DcTerms {
  const float* y = nullptr;
  const float* density = nullptr;  // [rows] or null (all ones)
  float scale = 1.0f;
  uint32_t row_stride = 0, coil_stride = 0;
};
// Forward interpolation of an interleaved split panel with the residual fused
// into its epilogue; writes the interleaved split residual panel.
// The dense-control tiles of a packed direction, expanded on the device from
// its sparse operands on first use (idempotent; not for concurrent callers).
int ensure_dense_tiles(const tsn_direction* d);
int launch_interp_dc(const tsn_direction* d, int kind, const uint32_t* panel, uint32_t coils, float scale,
                     const DcTerms& dc, uint32_t* out_il, hipStream_t s, const ChunkScratch* scratch = nullptr);
// The same residual from stored predictions float2 [rows][coils].
int launch_dc_residual(const float2* pred, uint32_t rows, uint32_t coils, const DcTerms& dc, bool bf16,
                       uint32_t* out_il, hipStream_t s);
// In place on FP32 rows float2 [rows][coils] (FP32 plans): pred <- residual.
int launch_dc_residual_rows(float2* rows_io, uint32_t rows, uint32_t coils, const DcTerms& dc, hipStream_t s);

}  // namespace tsn

struct tsn_host_pack : tsn::HostPack {};

struct tsn_kb_matrix {
  uint32_t rows = 0, width = 0;          // samples, numpoints^2
  uint32_t image0 = 0, image1 = 0, grid0 = 0, grid1 = 0, numpoints = 0;
  std::vector<int32_t> cols;             // [rows * width] ascending per row
  std::vector<float> re, im;             // float32 components
  std::vector<double> coord0, coord1;    // mod(traj*grid + grid/2, grid), float64
};

struct tsn_direction {
  tsn_pack_header h{};
  int device = 0;
  bool csr = false;
  int values = TSN_VALUES_FP16;      // tile value format (TSN_VALUES_*)
  // ---- packed 2:4 tiles (device)
  uint32_t* group_offsets = nullptr;
  int32_t* group_rows = nullptr;     // [groups * 16] logical output rows, -1 = padding
  uint32_t* tile_cols = nullptr;     // [tiles * 32] column * 32 | last-tile flag (bit 31, word 2 per 8)
  uint4* tile_a = nullptr;           // [tiles * 64] per lane {re 4xf16, im 4xf16}
  uint16_t* tile_idx = nullptr;      // [tiles * 64] per lane sparse index byte re | im << 8
  uint4* tile_dense = nullptr;       // [tiles * 64 * 2] per lane {re 8xf16}, {im 8xf16}; built on first use
  tsn::WorkItem* items = nullptr;    // runs of groups with <= kSplitAboveTiles tiles each
  tsn::ChunkItem* chunks = nullptr;  // chunks of the longer groups (fixed by the pack)
  float* partials = nullptr;         // [slots][kSlotFloats] partial sums of the chunks
  uint32_t* counters = nullptr;      // [long groups] arrival counters, zero between launches
  uint32_t n_items = 0, n_chunks = 0;
  uint32_t n_slots = 0, n_counters = 0;  // sizes of the chunk scratch (ChunkScratch)
  uint32_t unit_tiles = 0;           // tile budget of one regular unit
  uint32_t coil_waves = tsn::kDefaultCoilWaves;  // interleaved launches of 9..16 coils: waves per entry (1 or 2)
  uint32_t pipe_depth = 0;  // tiles in flight ahead of the multiply: 0 = the kernel's default, 1..3
  std::vector<uint32_t> host_offsets;  // group_offsets kept for rescheduling
  uint32_t* order = nullptr;         // [n_items + n_chunks] launch order (kChunkTag marks chunks)
  std::vector<tsn::WorkItem> host_items;  // the current schedule (batched launches merge it)
  std::vector<uint32_t> host_order;
  std::vector<void*> retired;          // earlier schedules (graphs may reference them)
  // ---- FP32 CSR (device)
  uint32_t csr_rows = 0;
  uint32_t* row_ptr = nullptr;   // [csr_rows + 1]
  int32_t* row_ids = nullptr;    // [csr_rows] logical output row
  int32_t* cols = nullptr;       // [nnz]
  float2* vals = nullptr;        // [nnz]
  tsn::CsrItem* csr_items = nullptr;  // [n_csr_items] work items in launch order
  uint32_t n_csr_items = 0;
  uint32_t n_csr_active_items = 0;  // leading items holding a row with entries (the rest zero empty rows)
  // (partials / counters / n_slots / n_counters above hold the long-row chunk
  // scratch of a CSR direction: kMaxCoils float2 per slot)
  size_t bytes = 0;
};
