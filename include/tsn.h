/*
 * tsn.h - C ABI of the MI355X (gfx950) fixed-trajectory NUFFT / CG-SENSE path.
 *
 * Plain C: pointers, sizes and status codes only.  Device pointers are HIP
 * device addresses; `stream` is a hipStream_t passed as void* (0 = null
 * stream).  Every entry point returns TSN_OK or a status code mirroring the
 * reference's process exit codes (nufft_spmma_bench.cu:61-69, 92-96;
 * nufft_integrated_bench.cu:141-145, 189-194; nufft_fp16x2_cg_bench.cu:
 * 317-322, 418-421).  The library never calls exit(); tsn_last_error()
 * returns the message of the last failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   pack directory format  src/export_torchkbnufft_gpu_case.py:387-436 (writer)
 *                          experiments/production/src/nufft_integrated_bench.cu:135-226 (reader)
 *   FFT endpoint ABI       experiments/production/src/nufft_custom_fft_f30_api.h:6-34
 *   operator / solver      nufft_integrated_bench.cu:431-466 (launch_fused_interpolation[_residual]),
 *                          nufft_fp16x2_cg_bench.cu:203-268 (launch_adjoint_image, launch_normal, launch_cg)
 */
#ifndef TSN_H
#define TSN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 1/2/3 carry the reference binaries' exit-code meaning. */
enum {
  TSN_OK = 0,
  TSN_ERR_RUNTIME = 1,      /* HIP error, unreadable/short file, bad header (ref exit 1) */
  TSN_ERR_INCOMPATIBLE = 2, /* real/imag or operator shape mismatch (ref exit 2) */
  TSN_ERR_CG = 3,           /* CG non-positive curvature or non-finite output (ref exit 3) */
  TSN_ERR_ARGUMENT = 4      /* invalid argument or unsupported configuration */
};

/* Interpolation precision / execution variants. */
enum {
  TSN_PREC_SPLIT16 = 0,  /* FP16 hi + FP16 residual panels, 2:4 sparse MFMA, FP32 accumulate */
  TSN_PREC_DENSE16 = 1,  /* same panels and tiles on dense MFMA (equal-quality control) */
  TSN_PREC_FP32 = 2,     /* plain FP32 CSR gather with FP32 coefficients */
  TSN_PREC_HI16 = 3,     /* FP16 hi panel only on 2:4 sparse MFMA */
  TSN_PREC_SPLITBF16 = 4 /* BF16 hi + BF16 residual panels, BF16 tiles, 2:4 sparse MFMA */
};

/* Output layouts of the interpolation entry points. */
enum {
  TSN_OUT_COIL_MAJOR = 0, /* float2 [coil][row]   (reference output layout) */
  TSN_OUT_ROW_MAJOR = 1,  /* float2 [row][coil]   (coil-interleaved rows)  */
  TSN_OUT_SPLIT_IL = 2    /* interleaved split panel of the output rows (tsn_interp_il only) */
};

/* Pack header, 52 bytes little endian, struct "<Q11I"
 * (reference src/export_gpu_case.py:242-256, nufft_spmma_bench.cu:42-59). */
#pragma pack(push, 1)
typedef struct tsn_pack_header {
  uint64_t magic; /* 0x4E55464654323431 "NUFFT241" */
  uint32_t version; /* 2 */
  uint32_t n_channels; /* 16 */
  uint32_t output_size;
  uint32_t n_cols;
  uint32_t active_rows;
  uint32_t nnz;
  uint32_t groups;
  uint32_t tiles;
  uint32_t residual_nnz; /* must be 0 */
  uint32_t m_tile; /* 16 */
  uint32_t k_tile; /* 32 */
} tsn_pack_header;
#pragma pack(pop)

const char* tsn_last_error(void);
/* Library build identification (for logs). */
const char* tsn_version(void);

/* ------------------------------------------------------------------ packs */

/* One packed direction (real + imag components sharing one structure), in
 * host memory.  Replaces the reference's per-component pack directories. */
typedef struct tsn_host_pack tsn_host_pack;

/* Read and validate a real/imag directory pair (reference load_direction,
 * nufft_integrated_bench.cu:135-194: same header checks, identical structure,
 * residual_nnz == 0) plus bounds checks of every index array. */
int tsn_host_pack_read(const char* real_dir, const char* imag_dir, tsn_host_pack** out);
/* Write the pair in the reference directory format
 * (export_torchkbnufft_gpu_case.py:407-436; production-only file set). */
int tsn_host_pack_write(const tsn_host_pack* pack, const char* real_dir, const char* imag_dir);
/* Build a pack from caller arrays (copied and validated). */
int tsn_host_pack_create(const tsn_pack_header* header, const uint32_t* group_offsets,
                         const int32_t* group_row_ids, const int32_t* tile_col_ids,
                         const uint16_t* real_comp, const uint32_t* real_meta,
                         const uint16_t* imag_comp, const uint32_t* imag_meta,
                         tsn_host_pack** out);
int tsn_host_pack_header(const tsn_host_pack* pack, tsn_pack_header* out);
/* Borrowed pointers valid until tsn_host_pack_free. */
int tsn_host_pack_arrays(const tsn_host_pack* pack, const uint32_t** group_offsets,
                         const int32_t** group_row_ids, const int32_t** tile_col_ids,
                         const uint16_t** real_comp, const uint32_t** real_meta,
                         const uint16_t** imag_comp, const uint32_t** imag_meta);
void tsn_host_pack_free(tsn_host_pack* pack);

/* ------------------------------------------------------------- KB matrix */

/* Complex KB interpolation matrix G (samples x grid), float32 components,
 * numpoints^2 entries per row, row-sorted columns.  Follows torchkbnufft
 * calc_tensor_spmatrix as called at export_torchkbnufft_gpu_case.py:537-542.
 * traj: [samples][2] float32 in cycles/pixel, [-0.5, 0.5); traj[m][0] is the
 * slow image axis.  threads <= 0 uses all hardware threads. */
typedef struct tsn_kb_matrix tsn_kb_matrix;
int tsn_kb_build(const float* traj, uint32_t samples, uint32_t image0, uint32_t image1,
                 uint32_t grid0, uint32_t grid1, uint32_t numpoints, int threads,
                 tsn_kb_matrix** out);
int tsn_kb_arrays(const tsn_kb_matrix* kb, uint32_t* rows, uint32_t* width, const int32_t** cols,
                  const float** real, const float** imag);
/* Apodization (torchkbnufft scaling_coef, real): float32 [image0][image1]. */
int tsn_kb_apodization(uint32_t image0, uint32_t image1, uint32_t grid0, uint32_t grid1,
                       uint32_t numpoints, float* out);
/* Pack view 0 (G, rows = samples) or 1 (G_T, rows = active grid points)
 * with the production planner (policy 0 = morton, 1 = trajectory order, 2 = overlap),
 * the reference default seed being 20260823. */
int tsn_kb_pack(const tsn_kb_matrix* kb, int view, int policy, uint64_t seed, int threads,
                tsn_host_pack** out);
/* Tile value formats.  FP16 is the reference format; a BF16 pack has the same
 * layout and planner, its kept values and metadata chosen on BF16-rounded
 * coefficients (no counterpart in the reference; used by TSN_PREC_SPLITBF16). */
enum { TSN_VALUES_FP16 = 0, TSN_VALUES_BF16 = 1 };
int tsn_kb_pack_values(const tsn_kb_matrix* kb, int view, int policy, uint64_t seed, int threads,
                       int values, tsn_host_pack** out);
void tsn_kb_free(tsn_kb_matrix* kb);

/* ------------------------------------------------------- device directions */

typedef struct tsn_direction tsn_direction;
/* Upload a pack to the current device (tiles re-laid out per MFMA lane). */
int tsn_direction_upload(const tsn_host_pack* pack, tsn_direction** out);
/* Reference load_direction: read + validate + upload. */
int tsn_direction_load(const char* real_dir, const char* imag_dir, tsn_direction** out);
/* FP32 CSR direction of the plain path: view 0 = G, 1 = G_T (exact transpose). */
int tsn_direction_upload_csr(const tsn_kb_matrix* kb, int view, tsn_direction** out);
int tsn_direction_header(const tsn_direction* dir, tsn_pack_header* out);
/* Re-cut the work schedule of a packed direction: regular units are runs of
 * whole groups holding at most unit_tiles tiles each (0 = default: the tiles
 * over 12 waves per CU, 4..14 tiles; larger values act as 16, the most a wave
 * stages in LDS).  Earlier schedules stay valid for graphs
 * captured before the call.  Synchronizes the device; no counterpart in the
 * reference (its launch maps one group per warp), used by the autotuner. */
int tsn_direction_schedule(tsn_direction* dir, uint32_t unit_tiles);
/* Tile budget of the current schedule (0 for CSR directions). */
uint32_t tsn_direction_unit_tiles(const tsn_direction* dir);
/* Waves per work entry of interleaved-panel interpolation of 9..16 coils
 * (tsn_interp_il, plans, batches): 1 = one wave computes both coil blocks
 * from the shared tile operands (the default, measured faster at 768^2 / 16
 * coils), 2 = one wave per coil block on the same entry.  Results are
 * bitwise the same.  No counterpart in the reference (8 coils, one warp per
 * group). */
int tsn_direction_set_coil_waves(tsn_direction* dir, uint32_t waves);
uint32_t tsn_direction_coil_waves(const tsn_direction* dir);
/* Pipeline depth of interleaved-panel interpolation (tiles whose operand
 * loads are in flight ahead of the multiply): 0 = the kernel's default (2
 * for one coil block per wave, 1 for two), or 1..3; a depth the launch does
 * not offer (1 or 3 for one block per wave, 2 for two; plan-style outputs)
 * falls back to the default.  Results are bitwise the same.  No counterpart
 * in the reference. */
int tsn_direction_set_pipe_depth(tsn_direction* dir, uint32_t depth);
uint32_t tsn_direction_pipe_depth(const tsn_direction* dir);
/* Autotuner: time the interpolation of this direction for `coils` coils over
 * the candidate unit sizes {2,3,4,5,6,8,10,12,14,16} x pipeline depths
 * (default and alternatives), and for 9..16 coils over both coil-wave
 * settings, writing 64-byte rows per coil block as plans do (scratch panels
 * allocated and freed here, the stream synchronized; two passes, the faster
 * of each candidate counts), and keep the fastest schedule.  Results do not
 * depend on the schedule (bitwise).  No counterpart in the reference; run
 * once per trajectory and coil count. */
int tsn_direction_autotune(tsn_direction* dir, uint32_t coils, void* stream, uint32_t* best_unit_tiles,
                           float* best_us);
/* Device bytes the direction occupies (pack + index arrays, plus the
 * dense-control tiles once they have been expanded). */
size_t tsn_direction_bytes(const tsn_direction* dir);
/* Expand the dense-control tiles (16x32 FP16, 2 KiB per tile) from the
 * direction's sparse operands on the device.  Uploads leave them out; dense16
 * plans, batches and tsn_interp_dense expand them on first use.  Idempotent;
 * synchronous.  Counterpart of experiments/packing/prepare_dense_control.py:27-87
 * (decode_row_metadata / expand_component), which writes them offline. */
int tsn_direction_expand_dense(tsn_direction* dir);
void tsn_direction_free(tsn_direction* dir);

/* Concurrency of the direct interpolation calls below (tsn_interp,
 * tsn_interp_dense, tsn_interp_il, tsn_interp_dc): a long group (more than 16
 * tiles) is summed by several waves through partial slots and an arrival
 * counter that the direction holds, so launches of these calls on one
 * direction must not overlap (one stream at a time, or ordered by events).
 * Plans and batches hold their own slots and counters and may share a
 * direction across streams.
 *
 * Interpolation y = G x (conjugate = 0) or G^H-style conjugate combine (1)
 * of FP16 N16 panels ([row][16] halves: Re coils 0..7, Im coils 0..7, blocks
 * of 8 coils, block stride n_cols*16 halves).  panel_res may be NULL (hi
 * only).  Output float2 per layout; accumulate adds to it.  scale multiplies
 * the result.  Only the rows of the pack's groups are written; other output
 * rows (grid points no sample reaches, for G^T) keep their contents, as in
 * the reference, whose harness zeroes the grid once.  Equivalent to
 * launch_fused_interpolation followed by launch_fused_interpolation_residual
 * (nufft_integrated_bench.cu:431-466). */
int tsn_interp(const tsn_direction* dir, const void* panel_hi, const void* panel_res, void* out,
               uint32_t coils, int conjugate, int accumulate, float scale, int out_layout,
               void* stream);
/* tsn_interp without any allocation or direction-held scratch: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The long-group chunk partials and arrival counters live
 * in one static device array per GPU (6 MiB + 16 KiB, held until the process
 * ends): at most 1536 chunks (4 KiB each) and 4096 long groups per direction,
 * else TSN_ERR_ARGUMENT.  The direction must live on the current device.  The
 * counters are cleared by a fill kernel on `stream` at each call.  Every call
 * on one GPU shares that array, so calls that overlap on one GPU (other
 * streams, other threads) race; order them.  Same arguments, kernels and
 * results (bitwise) as tsn_interp. */
int tsn_interp_noalloc(const tsn_direction* dir, const void* panel_hi, const void* panel_res, void* out,
                       uint32_t coils, int conjugate, int accumulate, float scale, int out_layout,
                       void* stream);
/* Same operator on the dense-MFMA control (tiles expanded to 16x32 on first use). */
int tsn_interp_dense(const tsn_direction* dir, const void* panel_hi, const void* panel_res,
                     void* out, uint32_t coils, int conjugate, int accumulate, float scale,
                     int out_layout, void* stream);
/* Plain FP32 path: CSR direction, input float2 [row][coils], output per layout.
 * Without accumulate every output row is written, zeros where no entry
 * reaches, like the full sparse product of the reference's FP32 path
 * (torch.sparse.mm, experiments/production/reference_sparse_solver.py:118-129). */
int tsn_interp_f32(const tsn_direction* csr, const void* in, void* out, uint32_t coils,
                   int conjugate, int accumulate, float scale, int out_layout, void* stream);
/* hi = RN16(v), res = RN16(v - hi) of float2 [coil][rows] * scale into N16 panel
 * blocks (reference integrated_complex_to_split2_half_rows, :339-359). */
int tsn_split_panel(const void* in, uint32_t rows, uint32_t coils, float scale, void* hi, void* res,
                    void* stream);
/* Interleaved split panel: the layout plans use internally (no counterpart in
 * the reference, which keeps hi and residual in two panels).  uint32 per
 * channel = hi | residual << 16, blocks of [rows][16] for ceil(coils/8) coil
 * blocks; one 64-byte gather per column instead of two.  values selects FP16
 * halves (TSN_VALUES_FP16, rounded exactly as tsn_split_panel) or bfloat16
 * halves (TSN_VALUES_BF16, round to nearest even).  tsn_interp_il is
 * tsn_interp on such a panel with the direction's own value format (same
 * tiles, same sums, same output for FP16).  With out_layout TSN_OUT_SPLIT_IL
 * (accumulate 0) it writes its result as such a panel of its output rows
 * (int32 [blocks][rows][16], the direction's value format): the forward
 * interpolation handing the adjoint its input, as inside a plan. */
int tsn_split_panel_il(const void* in, uint32_t rows, uint32_t coils, float scale, int values,
                       void* panel, void* stream);
int tsn_interp_il(const tsn_direction* dir, const void* panel, void* out, uint32_t coils,
                  int conjugate, int accumulate, float scale, int out_layout, void* stream);

/* Data consistency of the forward direction (JKF_DATA_CONSISTENCY_MODE 1/2 of
 * nufft_integrated_bench.cu:67-82, 497-525, 700-715; epilogue
 * nufft_spmma_bench.cu:561-603 and integrated_residual_to_half_rows :375-408):
 *   r[row][c] = (scale * (G panel)[row][c] - measurement_scale * y[row][c]) * density[row]
 * written as an interleaved split panel of the direction's value format
 * (the adjoint's input, as tsn_split_panel_il lays it out).  y is float2 in
 * measurement_layout (TSN_OUT_COIL_MAJOR [coils][rows], the reference's, or
 * TSN_OUT_ROW_MAJOR [rows][coils], one 64-byte read per row); density float
 * [rows] or NULL for all ones.  TSN_DC_MATERIALIZED interpolates into
 * workspace (float2 [rows][coils]) and forms r in a second pass;
 * TSN_DC_FUSED forms r in the interpolation epilogue.  Both give the same
 * bits (each step rounds on its own). */
enum { TSN_DC_MATERIALIZED = 1, TSN_DC_FUSED = 2 };
int tsn_interp_dc(const tsn_direction* forward, const void* panel, const void* measurements,
                  int measurement_layout, const float* density, uint32_t coils, float scale,
                  float measurement_scale, int mode, void* workspace, void* out_panel, void* stream);

/* ------------------------------------------------------------------ plan */

typedef struct tsn_plan_config {
  uint32_t image0, image1; /* image shape (n0 slow) */
  uint32_t grid0, grid1;   /* oversampled grid = 2*image, sides in {64,128,256,512,768} */
  uint32_t coils;          /* 1..16 */
  int precision;           /* TSN_PREC_* */
} tsn_plan_config;

typedef struct tsn_plan tsn_plan;
/* forward: G pack (rows = samples), adjoint: G_T pack (rows = grid points);
 * both CSR directions for TSN_PREC_FP32.  smaps: device float2 [coils][image0][image1];
 * apodization: device float [image0][image1] (torchkbnufft scaling_coef).
 * The plan copies what it needs; all work buffers are allocated here. */
int tsn_plan_create(const tsn_plan_config* config, const tsn_direction* forward,
                    const tsn_direction* adjoint, const void* smaps, const void* apodization,
                    tsn_plan** out);
void tsn_plan_free(tsn_plan* plan);
size_t tsn_plan_bytes(const tsn_plan* plan);

/* A x: image float2 [image0][image1] -> samples float2 [coils][M], normalised by
 * 1/sqrt(grid0*grid1) (the native scaling of nufft_fp16x2_cg_bench.cu:230-236). */
int tsn_forward(tsn_plan* plan, const void* image, void* samples, void* stream);
/* A^H y (exact adjoint of tsn_forward). */
int tsn_adjoint(tsn_plan* plan, const void* samples, void* image, void* stream);
/* (A^H A) x * inv_scale_sq + scaled_lambda * x (launch_normal). */
int tsn_normal(tsn_plan* plan, const void* image, void* out, float inv_scale_sq,
               float scaled_lambda, void* stream);

/* Data-consistency gradient A^H W (A x - measurement_scale * y) in the
 * operator normalisation of tsn_forward/tsn_adjoint (pass measurement_scale =
 * 1/sqrt(grid0*grid1) for raw torchkbnufft-scale data): image and out float2
 * [image0][image1], y float2 [coils][M], density float [M] or NULL (ones).
 * Split-panel precisions form the weighted residual in the forward
 * interpolation epilogue (TSN_DC_FUSED); FP32 in a pass over the CSR output.
 * The step either side of the forward interpolation in unrolled and
 * regularised solvers (nufft_integrated_bench.cu:700-715 + the SENSE
 * combine of nufft_fp16x2_cg_bench.cu:27-54). */
int tsn_data_consistency(tsn_plan* plan, const void* image, const void* measurement, const float* density,
                         float measurement_scale, void* out, void* stream);

/* CG-SENSE (launch_cg): measurement raw float2 [coils][M] (torchkbnufft scale),
 * operator_scale as the reference CLI argument (native scale), lambda the raw
 * Tikhonov weight (reference 1e-4).  Writes x float2 [image0][image1], the
 * residual history sqrt(rho) as double[iterations+1] and a status int
 * (1 = ok, 0 = non-positive curvature) to device memory, asynchronously.
 * CG scalars (rho, <p,Ap>, alpha, beta) are float64 on the device. */
int tsn_cg(tsn_plan* plan, const void* measurement, uint32_t iterations, float operator_scale,
           float lambda, void* x_out, void* residuals_out, void* status_out, void* stream);
/* Capture the same solve into a hipGraph bound to these pointers; launch it
 * any number of times; results are bitwise equal to tsn_cg. */
int tsn_cg_graph_capture(tsn_plan* plan, const void* measurement, uint32_t iterations,
                         float operator_scale, float lambda, void* x_out, void* residuals_out,
                         void* status_out, void* stream);
int tsn_cg_graph_launch(tsn_plan* plan, void* stream);

/* ------------------------------------------------------------- batches */

/* Many independent CG-SENSE problems of one image / grid shape solved in
 * lockstep: every step is one launch for all of them (the FFT passes and the
 * vector kernels index the problem, the interpolation merges all problems'
 * work units; problems sharing a pack read its tiles once per step).  Each
 * problem's result equals its tsn_cg result bit for bit.  No counterpart in
 * the reference, which reconstructs one slice per process; for uneven slice
 * batches (spokes, coils) and frame series sharing one packed trajectory. */
typedef struct tsn_batch tsn_batch;
typedef struct tsn_batch_problem {
  const tsn_direction* forward; /* G pack (rows = samples) */
  const tsn_direction* adjoint; /* G_T pack */
  const void* smaps;            /* device float2 [coils][image0][image1] (copied) */
  uint32_t coils;               /* 1..16 */
} tsn_batch_problem;
/* cfg: image / grid / precision (split16, dense16, hi16 or splitbf16; coils
 * ignored).  apodization: device float [image0][image1], shared. */
int tsn_batch_create(const tsn_plan_config* cfg, const tsn_batch_problem* problems, uint32_t count,
                     const void* apodization, tsn_batch** out);
void tsn_batch_free(tsn_batch* batch);
size_t tsn_batch_bytes(const tsn_batch* batch);
/* Per problem z: measurements[z] raw float2 [coils][M_z], operator_scales[z],
 * x_out[z] float2 [image0][image1], residuals_out[z] double[iterations + 1];
 * status_out int[count] (1 ok, 0 non-positive curvature).  The host arrays are
 * read during the call (device tables are staged after draining the stream). */
/* Batch autotuner: time the batch's interpolation steps (every problem's
 * forward and adjoint launches) over the member group of frames sharing a
 * pack (how many of them follow each work entry: 1, 2, 4, 8, 16 or all), the
 * pipeline depth (0..3, as tsn_direction_set_pipe_depth, one value for all the
 * batch's launches) and, when problems of 9..16 coils are present, 1 or 2
 * waves per entry; keep the fastest (two passes, the faster of each candidate
 * counts).  Results are bitwise the same.  Call before capturing a graph
 * (fails otherwise); synchronizes the stream.  The batch's interpolation
 * buffers hold unspecified values afterwards (every solve rewrites them). */
int tsn_batch_autotune(tsn_batch* batch, void* stream, uint32_t* best_member_group, uint32_t* best_coil_waves,
                       uint32_t* best_depth, float* best_us);
int tsn_batch_cg(tsn_batch* batch, const void* const* measurements, uint32_t iterations,
                 const float* operator_scales, float lambda, void* const* x_out,
                 void* const* residuals_out, void* status_out, void* stream);
int tsn_batch_cg_graph_capture(tsn_batch* batch, const void* const* measurements, uint32_t iterations,
                               const float* operator_scales, float lambda, void* const* x_out,
                               void* const* residuals_out, void* status_out, void* stream);
int tsn_batch_cg_graph_launch(tsn_batch* batch, void* stream);

/* ----------------------------------------------- FFT endpoint drop-in ABI */
/* Same names and argument meaning as nufft_custom_fft_f30_api.h:6-34 (fixed
 * 256^2 image, 512^2 grid, 8 coils).  Pointers: image float2, scaling float,
 * transposed_workspace float2 [8][512][512] (pre-zeroed), N16 outputs __half
 * [512*512][16], stream hipStream_t.  Returns hipError_t as int. */
int jkf_initialize_custom_fft_twiddles(void);
void jkf_launch_custom_fft_f30_forward(const void* image, const float* scaling,
                                       void* transposed_workspace, void* n16_output, void* stream);
void jkf_launch_custom_fft_f30_forward_split2(const void* image, const float* scaling,
                                              void* transposed_workspace, void* n16_high_output,
                                              void* n16_residual_output, void* stream);
void jkf_launch_custom_fft_f30_forward_split2_debug(const void* image, const float* scaling,
                                                    void* transposed_workspace,
                                                    void* n16_high_output,
                                                    void* n16_residual_output,
                                                    void* fp32_endpoint_output, void* stream);
void jkf_launch_custom_fft_f30_forward_sense(const void* image, const void* sensitivity_maps,
                                             const float* scaling, void* transposed_workspace,
                                             void* n16_output, void* stream);
void jkf_launch_custom_fft_f30_forward_sense_split2(const void* image,
                                                    const void* sensitivity_maps,
                                                    const float* scaling,
                                                    void* transposed_workspace,
                                                    void* n16_high_output,
                                                    void* n16_residual_output, void* stream);
void jkf_launch_custom_ifft_f31_adjoint(const void* grid_input, void* transposed_workspace,
                                        const float* scaling, const float* scaling_tiled,
                                        void* image_output, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TSN_H */
