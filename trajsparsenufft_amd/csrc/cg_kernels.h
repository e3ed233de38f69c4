// Vector kernels shared by plans and batches: CG updates with float64
// scalars (blockIdx.y = problem of a batched launch, pointers advanced by the
// given strides; a single plan passes strides 0), transposes, fills.
#pragma once

#include <cmath>
#include <vector>

#include "fft_common.h"

namespace {

using tsn::div_up;
using tsn::fail;

constexpr int kVecThreads = 256;

// The last CG step's update (x taken as zero with x_zero: a one-step solve):
// alpha = rho / <p, Ap> from the inverse pass's
// partials dots [3][dot_count] (<p, Ap>, Re<r, Ap>, <Ap, Ap>; 0 and status 0
// on non-positive or non-finite curvature), x += alpha p; block 0 writes
// rho_hist[iter + 1] = ||r - alpha Ap||^2 (expanded as in the forward pass).
// Every block reads rho_hist[iter] and nothing else of the history, so the
// one write cannot race with a read; the residual history sqrt(rho) is formed
// by a later launch (batch_out_kernel).
__global__ __launch_bounds__(kVecThreads) void cg_final_kernel(
    float2* __restrict__ x, const float2* __restrict__ p, const double* __restrict__ dots, uint32_t dot_count,
    double* __restrict__ rho_hist, uint32_t iter, int x_zero, int* __restrict__ status, uint32_t n,
    uint64_t vstride, uint32_t dstride, uint32_t hstride) {
  __shared__ double red[kVecThreads];
  const uint64_t z = blockIdx.y;  // problem of a batched launch
  x += z * vstride;
  p += z * vstride;
  dots += z * dstride;
  rho_hist += z * hstride;
  status += z;
  double pap, rap, apap;
  tsn::sum_partials3<kVecThreads / 64>(dots, dot_count, pap, rap, apap, red);
  const double rho = rho_hist[iter];
  double alpha = 0.0;
  if (isfinite(pap) && pap > 0.0) alpha = rho / pap;
  else if (blockIdx.x == 0 && threadIdx.x == 0) *status = 0;
  const uint32_t e = blockIdx.x * kVecThreads + threadIdx.x;
  if (e < n) {
    const float2 pv = p[e];
    float2 xv = x_zero ? make_float2(0.f, 0.f) : x[e];
    xv.x = static_cast<float>(xv.x + alpha * pv.x);
    xv.y = static_cast<float>(xv.y + alpha * pv.y);
    x[e] = xv;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    rho_hist[iter + 1] = fmax(rho - 2.0 * alpha * rap + alpha * alpha * apap, 0.0);
}

// The last CG step fused with the result (single plans): alpha as in
// cg_final_kernel, then x_out = transpose(x + alpha p) (x taken as zero with
// x_zero: a one-step solve), [cols][rows] from the
// column-major [rows][cols] vectors, in 32 x 32 tiles through LDS (256
// threads: 32 columns x 8 rows).  Block (0, 0) also writes the residual
// history hist_out[i] = sqrt(rho_i), i <= iter + 1 (rho_{iter+1} expanded as
// in the forward pass), and copies the status word to status_out.  rho_hist
// is only read here, and hist_out is a separate buffer, so no block can see a
// value another block rewrote.
__global__ __launch_bounds__(kVecThreads) void cg_final_out_kernel(
    const float2* __restrict__ x, const float2* __restrict__ p, const double* __restrict__ dots,
    uint32_t dot_count, const double* __restrict__ rho_hist, uint32_t iter, int x_zero, int* __restrict__ status,
    int* __restrict__ status_out, uint32_t rows, uint32_t cols, float2* __restrict__ x_out,
    double* __restrict__ hist_out) {
  __shared__ double red[kVecThreads];
  __shared__ float2 tile[32][33];
  const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const uint32_t bx = blockIdx.x * 32, by = blockIdx.y * 32;
  // issue the tile loads before the reductions so their latency overlaps
  float2 xv[4], pv[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t r = by + ty + 8 * k, c = bx + tx;
    const bool in = r < rows && c < cols;
    xv[k] = in && !x_zero ? x[static_cast<size_t>(r) * cols + c] : make_float2(0.f, 0.f);
    pv[k] = in ? p[static_cast<size_t>(r) * cols + c] : make_float2(0.f, 0.f);
  }
  double pap, rap, apap;
  tsn::sum_partials3<kVecThreads / 64>(dots, dot_count, pap, rap, apap, red);
  const double rho = rho_hist[iter];
  const bool first = blockIdx.x == 0 && blockIdx.y == 0;
  double alpha = 0.0;
  const bool curved = isfinite(pap) && pap > 0.0;
  if (curved) alpha = rho / pap;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    tile[ty + 8 * k][tx] = make_float2(static_cast<float>(xv[k].x + alpha * pv[k].x),
                                       static_cast<float>(xv[k].y + alpha * pv[k].y));
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t c = bx + ty + 8 * k, r = by + tx;
    if (r < rows && c < cols) x_out[static_cast<size_t>(c) * rows + r] = tile[tx][ty + 8 * k];
  }
  if (!first) return;
  for (uint32_t i = threadIdx.x; i <= iter; i += blockDim.x) hist_out[i] = sqrt(fmax(rho_hist[i], 0.0));
  if (threadIdx.x == 0) {
    hist_out[iter + 1] = sqrt(fmax(rho - 2.0 * alpha * rap + alpha * alpha * apap, 0.0));
    if (!curved) *status = 0;
    *status_out = *status;
  }
}

// [rows][cols] -> [cols][rows] complex, 32x32 tiles, optional scale.
__global__ __launch_bounds__(256) void transpose_kernel(const float2* __restrict__ in, uint32_t rows,
                                                        uint32_t cols, float scale,
                                                        float2* __restrict__ out) {
  __shared__ float2 tile[32][33];
  const uint32_t bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (uint32_t i = threadIdx.y; i < 32; i += 8) {
    const uint32_t r = by + i, c = bx + threadIdx.x;
    if (r < rows && c < cols) tile[i][threadIdx.x] = in[static_cast<size_t>(r) * cols + c];
  }
  __syncthreads();
  for (uint32_t i = threadIdx.y; i < 32; i += 8) {
    const uint32_t c = bx + i, r = by + threadIdx.x;
    if (r < rows && c < cols) {
      const float2 v = tile[threadIdx.x][i];
      out[static_cast<size_t>(c) * rows + r] = make_float2(v.x * scale, v.y * scale);
    }
  }
}

// [coils][n0][n1] -> [n1][cstride][n0]
__global__ __launch_bounds__(256) void smaps_transpose_kernel(const float2* __restrict__ in, uint32_t coils,
                                                              uint32_t n0, uint32_t n1, uint32_t cstride,
                                                              float2* __restrict__ out) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t total = static_cast<uint64_t>(coils) * n0 * n1;
  if (e >= total) return;
  const uint32_t x1 = static_cast<uint32_t>(e % n1);
  const uint32_t x0 = static_cast<uint32_t>((e / n1) % n0);
  const uint32_t c = static_cast<uint32_t>(e / (static_cast<uint64_t>(n1) * n0));
  out[(static_cast<size_t>(x1) * cstride + c) * n0 + x0] = in[e];
}

__global__ __launch_bounds__(256) void apod_transpose_kernel(const float* __restrict__ in, uint32_t n0,
                                                             uint32_t n1, float* __restrict__ out) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<uint64_t>(n0) * n1) return;
  const uint32_t x1 = static_cast<uint32_t>(e % n1), x0 = static_cast<uint32_t>(e / n1);
  out[static_cast<size_t>(x1) * n0 + x0] = in[e];
}

__global__ __launch_bounds__(256) void fill_zero_kernel(float4* __restrict__ p, uint64_t n4) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n4) p[e] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// [coils][m] -> [m][coils] * scale (FP32 path sample layout)
__global__ __launch_bounds__(256) void samples_to_rows_kernel(const float2* __restrict__ in, uint32_t coils,
                                                              uint32_t m, float scale,
                                                              float2* __restrict__ out) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<uint64_t>(coils) * m) return;
  const uint32_t c = static_cast<uint32_t>(e % coils), row = static_cast<uint32_t>(e / coils);
  const float2 v = in[static_cast<size_t>(c) * m + row];
  out[e] = make_float2(v.x * scale, v.y * scale);
}

// a fill kernel (not hipMemsetAsync) so the work can sit inside a captured graph
inline int zero_async(void* p, size_t bytes, hipStream_t s) {
  if (bytes % 16) return fail(TSN_ERR_ARGUMENT, "zero fill needs a multiple of 16 bytes");
  const uint64_t n4 = bytes / 16;
  if (n4 == 0) return TSN_OK;
  hipLaunchKernelGGL(fill_zero_kernel, dim3(div_up(n4, 256)), dim3(256), 0, s, static_cast<float4*>(p), n4);
  TSN_HIP(hipGetLastError());
  return TSN_OK;
}

// W_K^k = exp(-2 pi i k / K), k < K, computed in double
inline std::vector<float2> twiddles(uint32_t k) {
  std::vector<float2> t(k);
  for (uint32_t i = 0; i < k; ++i) {
    const double ang = -2.0 * M_PI * static_cast<double>(i) / static_cast<double>(k);
    t[i] = make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang)));
  }
  return t;
}

}  // namespace
