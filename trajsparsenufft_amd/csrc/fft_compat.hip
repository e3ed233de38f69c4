// FFT endpoints behind the reference's C ABI (nufft_custom_fft_f30_api.h:6-34),
// fixed to its production shape: 256^2 images, 512^2 grid, 8 coils.
//
// Layouts are the reference's: coil images float2 [coil][y][x], scaling float
// [y][x], transposed workspace float2 [coil][512][512] (ws[c][kx][y], rows
// y >= 256 never written), N16 panels __half [ky*512 + kx][16] (channel c = Re
// coil c, c + 8 = Im coil c), FP32 debug endpoint float2 [coil][ky][kx], grid
// input float2 [coil][ky][kx].  The inverse computes
// image = crop(IDFT2(grid)) * scaling / 512^2 (reference
// ifft512_rows8_regular_crop_scale, :1517-1552); `scaling_tiled` is accepted
// and unused, as in the reference's production inverse path.
//
// Each launch is two wave64-FFT passes built from fft_common.h: eight waves of
// a block transform eight rows (or columns) of one coil and exchange through
// LDS so every global access is a 64-byte or longer contiguous run.
#include <mutex>

#include "fft_common.h"

namespace tsn {
namespace {

constexpr int kR = 8;       // 512 = 64 * 8
constexpr int kG = 512;     // grid side
constexpr int kI = 256;     // image side
constexpr int kC = 8;       // coils

__device__ float2 g_twiddle512[kG];

__device__ __forceinline__ size_t lds_elems() { return 8 * wave_lds_elems<kR>(); }

// rows y0..y0+7 of coil c: x * [S_c] * scaling along x -> ws[c][kx][y]
template <bool SENSE>
__global__ __launch_bounds__(512) void compat_fwd_rows(const float2* __restrict__ image,
                                                       const float2* __restrict__ smaps,
                                                       const float* __restrict__ scaling,
                                                       float2* __restrict__ ws) {
  extern __shared__ float2 lds[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const WaveTwiddles<kR> twp = load_twiddles<kR>(g_twiddle512, lane);
  const int c = blockIdx.y, y = blockIdx.x * 8 + w;
  float2 x[kR];
#pragma unroll
  for (int j = 0; j < kR; ++j) {
    x[j] = make_float2(0.f, 0.f);
    if (j < kR / 2) {
      const int xi = lane + 64 * j;
      const size_t pix = static_cast<size_t>(y) * kI + xi;
      float2 v = SENSE ? cmul(image[pix], smaps[static_cast<size_t>(c) * kI * kI + pix])
                       : image[static_cast<size_t>(c) * kI * kI + pix];
      const float s = scaling[pix];
      x[j] = make_float2(v.x * s, v.y * s);
    }
  }
  float2 out[1][8];
  wave_fft<kR, false, true, false>(x, lds + w * wave_lds_elems<kR>(), g_twiddle512, twp, lane, out);
  __syncthreads();
  emit_exchange<kR, false>(out, lds, lane, w);  // element row of exchange row kx
  __syncthreads();
  const int kx = threadIdx.x;
  float4* dst = reinterpret_cast<float4*>(ws + (static_cast<size_t>(c) * kG + kx) * kG + blockIdx.x * 8);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float2 u = lds[xslot(kx, 2 * i)], v = lds[xslot(kx, 2 * i + 1)];
    dst[i] = make_float4(u.x, u.y, v.x, v.y);
  }
}

// column kx of all coils along y -> N16 hi [+ residual] panel [+ fp32 debug]
__global__ __launch_bounds__(512) void compat_fwd_cols(const float2* __restrict__ ws,
                                                       _Float16* __restrict__ hi,
                                                       _Float16* __restrict__ res,
                                                       float2* __restrict__ debug) {
  extern __shared__ float2 lds[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const WaveTwiddles<kR> twp = load_twiddles<kR>(g_twiddle512, lane);
  const int kx = blockIdx.x, c = w;
  float2 x[kR];
  const float2* col = ws + (static_cast<size_t>(c) * kG + kx) * kG;
#pragma unroll
  for (int j = 0; j < kR; ++j) x[j] = j < kR / 2 ? col[lane + 64 * j] : make_float2(0.f, 0.f);
  float2 out[1][8];
  wave_fft<kR, false, true, false>(x, lds + w * wave_lds_elems<kR>(), g_twiddle512, twp, lane, out);
  __syncthreads();
  emit_exchange<kR, false>(out, lds, lane, w);  // element c of exchange row ky
  __syncthreads();
  const int ky = threadIdx.x;
  const size_t r = static_cast<size_t>(ky) * kG + kx;
  half8 hre, him, rre, rim;
#pragma unroll
  for (int cc = 0; cc < kC; ++cc) {
    const float2 v = lds[xslot(ky, cc)];
    _Float16 h, q;
    split16(v.x, h, q);
    hre[cc] = h;
    rre[cc] = q;
    split16(v.y, h, q);
    him[cc] = h;
    rim[cc] = q;
    if (debug) debug[(static_cast<size_t>(cc) * kG + ky) * kG + kx] = v;
  }
  uint4* dh = reinterpret_cast<uint4*>(hi + r * 16);
  dh[0] = __builtin_bit_cast(uint4, hre);
  dh[1] = __builtin_bit_cast(uint4, him);
  if (res) {
    uint4* dr = reinterpret_cast<uint4*>(res + r * 16);
    dr[0] = __builtin_bit_cast(uint4, rre);
    dr[1] = __builtin_bit_cast(uint4, rim);
  }
}

// rows ky0..ky0+7 of coil c along kx, keep x < 256 -> ws[c][x][ky]
__global__ __launch_bounds__(512) void compat_inv_rows(const float2* __restrict__ grid,
                                                       float2* __restrict__ ws) {
  extern __shared__ float2 lds[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const WaveTwiddles<kR> twp = load_twiddles<kR>(g_twiddle512, lane);
  const int c = blockIdx.y, ky = blockIdx.x * 8 + w;
  const float2* row = grid + (static_cast<size_t>(c) * kG + ky) * kG;
  float2 x[kR];
#pragma unroll
  for (int j = 0; j < kR; ++j) x[j] = row[lane + 64 * j];
  float2 out[1][8];
  wave_fft<kR, true, false, true>(x, lds + w * wave_lds_elems<kR>(), g_twiddle512, twp, lane, out);
  __syncthreads();
  emit_exchange<kR, true>(out, lds, lane, w);  // element row of exchange row x
  __syncthreads();
  if (threadIdx.x < kI) {
    const int xi = threadIdx.x;
    float4* dst = reinterpret_cast<float4*>(ws + (static_cast<size_t>(c) * kG + xi) * kG + blockIdx.x * 8);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float2 u = lds[xslot(xi, 2 * i)], v = lds[xslot(xi, 2 * i + 1)];
      dst[i] = make_float4(u.x, u.y, v.x, v.y);
    }
  }
}

// columns x0..x0+7 of coil c along ky, keep y < 256, * scaling / 512^2
__global__ __launch_bounds__(512) void compat_inv_cols(const float2* __restrict__ ws,
                                                       const float* __restrict__ scaling,
                                                       float2* __restrict__ image) {
  extern __shared__ float2 lds[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const WaveTwiddles<kR> twp = load_twiddles<kR>(g_twiddle512, lane);
  const int c = blockIdx.y, x0 = blockIdx.x * 8;
  const float2* col = ws + (static_cast<size_t>(c) * kG + x0 + w) * kG;
  float2 x[kR];
#pragma unroll
  for (int j = 0; j < kR; ++j) x[j] = col[lane + 64 * j];
  float2 out[1][8];
  wave_fft<kR, true, false, true>(x, lds + w * wave_lds_elems<kR>(), g_twiddle512, twp, lane, out);
  __syncthreads();
  emit_exchange<kR, true>(out, lds, lane, w);  // element col of exchange row y
  __syncthreads();
  for (int e = threadIdx.x; e < kI * 8; e += blockDim.x) {
    const int y = e >> 3, xi = e & 7;
    const float s = scaling[y * kI + x0 + xi] * (1.0f / static_cast<float>(kG * kG));
    const float2 v = lds[xslot(y, xi)];
    image[(static_cast<size_t>(c) * kI + y) * kI + x0 + xi] = make_float2(v.x * s, v.y * s);
  }
}

std::mutex g_twiddle_mutex;

}  // namespace
}  // namespace tsn

using namespace tsn;

extern "C" {

int jkf_initialize_custom_fft_twiddles(void) {
  std::lock_guard<std::mutex> lock(g_twiddle_mutex);
  float2 tw[kG];
  for (int i = 0; i < kG; ++i) {
    const double ang = -2.0 * M_PI * i / kG;
    tw[i] = make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang)));
  }
  return static_cast<int>(hipMemcpyToSymbol(HIP_SYMBOL(g_twiddle512), tw, sizeof(tw)));
}

static void forward_rows(const void* image, const void* smaps, const float* scaling, void* ws, hipStream_t s) {
  const size_t lds = 8 * wave_lds_elems<kR>() * sizeof(float2);
  if (smaps)
    hipLaunchKernelGGL(compat_fwd_rows<true>, dim3(kI / 8, kC), dim3(512), lds, s,
                       static_cast<const float2*>(image), static_cast<const float2*>(smaps), scaling,
                       static_cast<float2*>(ws));
  else
    hipLaunchKernelGGL(compat_fwd_rows<false>, dim3(kI / 8, kC), dim3(512), lds, s,
                       static_cast<const float2*>(image), nullptr, scaling, static_cast<float2*>(ws));
}

static void forward_cols(const void* ws, void* hi, void* res, void* debug, hipStream_t s) {
  const size_t lds = 8 * wave_lds_elems<kR>() * sizeof(float2);
  hipLaunchKernelGGL(compat_fwd_cols, dim3(kG), dim3(512), lds, s, static_cast<const float2*>(ws),
                     static_cast<_Float16*>(hi), static_cast<_Float16*>(res), static_cast<float2*>(debug));
}

void jkf_launch_custom_fft_f30_forward(const void* image, const float* scaling, void* ws, void* n16,
                                       void* stream) {
  const hipStream_t s = static_cast<hipStream_t>(stream);
  forward_rows(image, nullptr, scaling, ws, s);
  forward_cols(ws, n16, nullptr, nullptr, s);
}

void jkf_launch_custom_fft_f30_forward_split2(const void* image, const float* scaling, void* ws, void* hi,
                                              void* res, void* stream) {
  const hipStream_t s = static_cast<hipStream_t>(stream);
  forward_rows(image, nullptr, scaling, ws, s);
  forward_cols(ws, hi, res, nullptr, s);
}

void jkf_launch_custom_fft_f30_forward_split2_debug(const void* image, const float* scaling, void* ws,
                                                    void* hi, void* res, void* fp32_out, void* stream) {
  const hipStream_t s = static_cast<hipStream_t>(stream);
  forward_rows(image, nullptr, scaling, ws, s);
  forward_cols(ws, hi, res, fp32_out, s);
}

void jkf_launch_custom_fft_f30_forward_sense(const void* image, const void* smaps, const float* scaling,
                                             void* ws, void* n16, void* stream) {
  const hipStream_t s = static_cast<hipStream_t>(stream);
  forward_rows(image, smaps, scaling, ws, s);
  forward_cols(ws, n16, nullptr, nullptr, s);
}

void jkf_launch_custom_fft_f30_forward_sense_split2(const void* image, const void* smaps, const float* scaling,
                                                    void* ws, void* hi, void* res, void* stream) {
  const hipStream_t s = static_cast<hipStream_t>(stream);
  forward_rows(image, smaps, scaling, ws, s);
  forward_cols(ws, hi, res, nullptr, s);
}

void jkf_launch_custom_ifft_f31_adjoint(const void* grid_input, void* ws, const float* scaling,
                                        const float* scaling_tiled, void* image_output, void* stream) {
  (void)scaling_tiled;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t lds = 8 * wave_lds_elems<kR>() * sizeof(float2);
  hipLaunchKernelGGL(compat_inv_rows, dim3(kG / 8, kC), dim3(512), lds, s,
                     static_cast<const float2*>(grid_input), static_cast<float2*>(ws));
  hipLaunchKernelGGL(compat_inv_cols, dim3(kI / 8, kC), dim3(512), lds, s, static_cast<const float2*>(ws),
                     scaling, static_cast<float2*>(image_output));
}

}  // extern "C"
