/*
 * Multi-core complex64 CSR product of the CPU baseline (oracle).
 *
 * TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times it, tests check
 * it; nothing in the product links it.
 *
 * out[r][c] = sum_k A[r][k] * x[col_k][c] for a complex64 CSR matrix A and a
 * complex64 panel x of `coils` columns (row-major [n_cols][coils]), FP32
 * arithmetic in entry order: the plain FP32 matrix x panel product of the
 * reference's fp32 emulation (experiments/production/reference_sparse_solver.py:
 * 131-197, torch.sparse.mm of the complex matrix on the stacked coil panel).
 * Rows are split over OpenMP threads (static schedule), so the result does not
 * depend on the thread count.
 */
#include <stdint.h>

void tsn_oracle_csr_c64(int64_t rows, const int64_t* row_ptr, const int32_t* cols, const float* vals,
                        const float* x, int coils, float* out, int threads) {
  if (threads < 1) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int64_t r = 0; r < rows; ++r) {
    float acc[2 * 64];
    for (int c = 0; c < 2 * coils; ++c) acc[c] = 0.0f;
    for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
      const float ar = vals[2 * k], ai = vals[2 * k + 1];
      const float* xr = x + (int64_t)cols[k] * 2 * coils;
      for (int c = 0; c < coils; ++c) {
        const float br = xr[2 * c], bi = xr[2 * c + 1];
        acc[2 * c] += ar * br - ai * bi;
        acc[2 * c + 1] += ar * bi + ai * br;
      }
    }
    float* o = out + r * 2 * coils;
    for (int c = 0; c < 2 * coils; ++c) o[c] = acc[c];
  }
}
