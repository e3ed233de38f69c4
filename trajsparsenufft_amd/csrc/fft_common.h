// Device building blocks of the wave64 FFT (shared by the plan kernels and
// the reference-ABI endpoints).
#pragma once

#include "tsn_internal.h"

namespace tsn {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));

struct FftArgs {
  uint32_t n0, n1;         // image shape
  uint32_t k0, k1;         // grid shape
  uint32_t coils;          // coils of the problem (per problem: coils_of[z] when set)
  uint32_t cstride;        // >= every problem's coils: SENSE-map channel stride, coil chunks of the
                           // launch, row width of the legacy grid layout
  const float2* tw0;       // W_K0^k, k < K0
  const float2* tw1;       // W_K1^k
  const float2* image_t;   // [n1][n0]
  const float2* smaps_t;   // [n1][coil][n0]
  const float* apod_t;     // [n1][n0]
  // Channel layouts (float2; see chunk_rows): t_fwd, t_inv and, with
  // grid_blocked, grid and panel32 hold each problem's rows compact (its
  // coils, up to 8) or in 8-coil chunk blocks of 64-byte rows; without
  // grid_blocked, grid and panel32 rows are cstride wide.
  float2* t_fwd;           // rows [K0][n1]
  uint32_t* panel_il;      // [block][K0*K1][16] hi | res << 16
  float2* panel32;         // rows [K0*K1]
  const float2* grid;      // rows [K0*K1]
  float2* t_inv;           // rows [n1][K0]
  uint32_t grid_blocked;
  float2* image_out;       // [n1][n0]
  const float2* reg_in;    // [n1][n0] or null
  float lambda;
  float out_scale;
  double* dot_partial;     // [n1] partials of Re<reg_in, out>, or null
  const float2* dot_r;     // with dot_partial: also Re<dot_r, out> at [n1 + n1] and <out, out> at [2 n1 + n1]
  // The previous CG step's update fused into the forward first axis (cg_r
  // null: the input is image_t).  Every block reduces the partials cg_dots
  // [3][dot_count] (<p, Ap>, Re<r, Ap>, <Ap, Ap>) in the same order: alpha =
  // rho_old / <p, Ap> (0 and *cg_status = 0 on non-positive or non-finite
  // curvature), rho_new = rho_old - 2 alpha Re<r, Ap> + alpha^2 <Ap, Ap>
  // (= ||r - alpha Ap||^2), beta = rho_new / max(rho_old, 1e-30); then
  // x += alpha p_old, r_new = r - alpha Ap, and the input is p_new = r_new +
  // beta p_old.  rho_hist[iter - 1] is read and rho_hist[iter] written by
  // block (0, 0); coil chunk 0's first wave stores x (in place: it alone reads
  // it), r_new and p_new (other buffers than r and p_old, which every chunk
  // and wave of the column reads).
  const float2* cg_r;
  float2* cg_r_new;
  float2* cg_x;
  const float2* cg_ap;
  const float2* cg_p_old;
  float2* cg_p_new;
  const double* cg_dots;
  uint32_t dot_count;
  int* cg_status;
  double* rho_hist;
  uint32_t iter;
  uint32_t x_zero;           // x is still zero (the first update): x = alpha p_old, x not read
  // First forward pass of a solve (cg_r null): rho0_dots [dot_count] holds the
  // partials of <b, b>; block (0, 0) writes their fixed-order sum to
  // rho_hist[0] and 1 to *cg_status.  Null: nothing.
  const double* rho0_dots;
  // Batched launches (blockIdx.z = problem z): every per-problem buffer sits
  // z strides (elements) from the pointers above; apod_t and the twiddles are
  // shared.  A single problem leaves them 0.
  const uint32_t* coils_of;  // [problems] or null
  const float2* scale_of;    // [problems] {out_scale, lambda} or null (inverse pass)
  uint64_t image_stride;     // image_t, image_out, reg_in, dot_r, cg_r, cg_r_new, cg_x, cg_ap, cg_p_old, cg_p_new
  uint64_t smaps_stride;     // smaps_t
  uint64_t t_stride;         // t_fwd, t_inv
  uint64_t grid_stride;      // grid, panel32
  uint64_t panel_stride;     // panel_il (uint32 words)
  uint64_t partial_stride;   // dot_partial, cg_dots, rho0_dots
  uint64_t hist_stride;      // rho_hist (cg_status: one int per problem)
};

// Sums of three values over a block of WAVES waves, in a fixed order, every
// thread getting all three: a butterfly over each wave (the two operands of
// every add are the same in all lanes, so all lanes hold the same bits), then
// the wave sums added in wave order.  One barrier pair; red holds 3 * WAVES
// doubles.  Every block running it on the same inputs gets the same bits.
template <int WAVES>
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
    c += __shfl_xor(c, off);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = a;
    red[WAVES + w] = b;
    red[2 * WAVES + w] = c;
  }
  __syncthreads();
  a = red[0];
  b = red[WAVES];
  c = red[2 * WAVES];
#pragma unroll
  for (int i = 1; i < WAVES; ++i) {
    a += red[i];
    b += red[WAVES + i];
    c += red[2 * WAVES + i];
  }
  __syncthreads();
}

// Fixed-order sums of the three partial arrays p[0 .. count), p[count ..
// 2 count), p[2 count .. 3 count) over a block of WAVES waves (1-D), every
// load issued before the first add.
template <int WAVES>
__device__ __forceinline__ void sum_partials3(const double* __restrict__ p, uint32_t count, double& a, double& b,
                                              double& c, double* red) {
  a = b = c = 0.0;
  for (uint32_t i = threadIdx.x; i < count; i += 64 * WAVES) {
    const double x = p[i], y = p[count + i], z = p[2 * count + i];
    a += x;
    b += y;
    c += z;
  }
  block_sum3<WAVES>(a, b, c, red);
}

// Where coil chunk `chunk` of row r starts (float2): chunk * cs + r * rs.  A
// problem of up to 8 coils keeps rows of exactly its coils (one chunk); above
// 8, every chunk is a block of rows of 8 (64 bytes, whole L2 sectors: a
// chunk never reads or writes part of a line that another chunk's block
// touches).  The legacy layout (grid and FP32 panel of plans whose grid the
// FP32 CSR kernel reads or writes) has rows of cstride with chunk c at
// column 8c.
struct ChunkRows {
  uint32_t rs;  // between rows
  uint64_t cs;  // between chunks
};
__device__ __forceinline__ ChunkRows chunk_rows(uint32_t coils, uint64_t rows) {
  return coils <= 8 ? ChunkRows{coils, 0} : ChunkRows{8u, rows * 8};
}

// The arguments of problem z of a batched launch.
__device__ __forceinline__ FftArgs problem_view(const FftArgs& a, uint32_t z) {
  FftArgs v = a;
  if (a.coils_of) v.coils = a.coils_of[z];
  if (a.scale_of) {
    v.out_scale = a.scale_of[z].x;
    v.lambda = a.scale_of[z].y;
  }
  if (z == 0) return v;
  const uint64_t zi = z;
  v.image_t = a.image_t ? a.image_t + zi * a.image_stride : nullptr;
  v.image_out = a.image_out ? a.image_out + zi * a.image_stride : nullptr;
  v.reg_in = a.reg_in ? a.reg_in + zi * a.image_stride : nullptr;
  v.dot_r = a.dot_r ? a.dot_r + zi * a.image_stride : nullptr;
  v.cg_r = a.cg_r ? a.cg_r + zi * a.image_stride : nullptr;
  v.cg_r_new = a.cg_r_new ? a.cg_r_new + zi * a.image_stride : nullptr;
  v.cg_x = a.cg_x ? a.cg_x + zi * a.image_stride : nullptr;
  v.cg_ap = a.cg_ap ? a.cg_ap + zi * a.image_stride : nullptr;
  v.cg_p_old = a.cg_p_old ? a.cg_p_old + zi * a.image_stride : nullptr;
  v.cg_p_new = a.cg_p_new ? a.cg_p_new + zi * a.image_stride : nullptr;
  v.smaps_t = a.smaps_t + zi * a.smaps_stride;
  v.t_fwd = a.t_fwd ? a.t_fwd + zi * a.t_stride : nullptr;
  v.t_inv = a.t_inv ? a.t_inv + zi * a.t_stride : nullptr;
  v.grid = a.grid ? a.grid + zi * a.grid_stride : nullptr;
  v.panel32 = a.panel32 ? a.panel32 + zi * a.grid_stride : nullptr;
  v.panel_il = a.panel_il ? a.panel_il + zi * a.panel_stride : nullptr;
  v.dot_partial = a.dot_partial ? a.dot_partial + zi * a.partial_stride : nullptr;
  v.cg_dots = a.cg_dots ? a.cg_dots + zi * a.partial_stride : nullptr;
  v.cg_status = a.cg_status ? a.cg_status + zi : nullptr;
  v.rho0_dots = a.rho0_dots ? a.rho0_dots + zi * a.partial_stride : nullptr;
  v.rho_hist = a.rho_hist ? a.rho_hist + zi * a.hist_stride : nullptr;
  return v;
}

// bfloat16 hi | residual << 16 of v (both round to nearest even).
__device__ __forceinline__ uint32_t split_bf16(float v) {
  const uint16_t h = bf16_rne_bits(v);
  return static_cast<uint32_t>(h) | (static_cast<uint32_t>(bf16_rne_bits(v - bf16_to_float(h))) << 16);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmul_conj_b(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// twiddle multiply: forward W, inverse conj(W)
template <bool INV>
__device__ __forceinline__ float2 twiddle(float2 a, float2 w) {
  return INV ? cmul_conj_b(a, w) : cmul(a, w);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

__device__ __forceinline__ void split16(float v, _Float16& hi, _Float16& res) {
  hi = static_cast<_Float16>(v);
  res = static_cast<_Float16>(v - static_cast<float>(hi));
}

// split16 of two values as interleaved words hi | res << 16, with one packed
// FP16 conversion per pair (the same round-to-nearest-even as split16): wa
// for a, wb for b.
__device__ __forceinline__ void split16_pair(float a, float b, uint32_t& wa, uint32_t& wb) {
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
  const f32x2 v = {a, b};
  const f16x2 hi = __builtin_convertvector(v, f16x2);
  const f32x2 back = __builtin_convertvector(hi, f32x2);
  const f32x2 rest = {v.x - back.x, v.y - back.y};
  const f16x2 res = __builtin_convertvector(rest, f16x2);
  const uint32_t hb = __builtin_bit_cast(uint32_t, hi), rb = __builtin_bit_cast(uint32_t, res);
  wa = __builtin_amdgcn_perm(rb, hb, 0x05040100u);  // bytes 0-1 of hb, then bytes 0-1 of rb
  wb = __builtin_amdgcn_perm(rb, hb, 0x07060302u);  // bytes 2-3 of hb, then bytes 2-3 of rb
}

// 2-point DFT; NZ = number of leading nonzero inputs (1 or 2).
template <bool INV, int NZ>
__device__ __forceinline__ void dft2(const float2* a, float2* X) {
  if constexpr (NZ == 2) {
    X[0] = cadd(a[0], a[1]);
    X[1] = csub(a[0], a[1]);
  } else {
    X[0] = a[0];
    X[1] = a[0];
  }
}

// 4-point DFT; NZ = number of leading nonzero inputs (2 or 4).
template <bool INV, int NZ>
__device__ __forceinline__ void dft4(const float2* a, float2* X) {
  float2 t0, t1, t2, t3;
  if constexpr (NZ == 4) {
    t0 = cadd(a[0], a[2]);
    t1 = csub(a[0], a[2]);
    t2 = cadd(a[1], a[3]);
    t3 = mul_mi<INV>(csub(a[1], a[3]));
  } else {
    t0 = a[0];
    t1 = a[0];
    t2 = a[1];
    t3 = mul_mi<INV>(a[1]);
  }
  X[0] = cadd(t0, t2);
  X[2] = csub(t0, t2);
  X[1] = cadd(t1, t3);
  X[3] = csub(t1, t3);
}

// 8-point DFT; NZ = number of leading nonzero inputs (4 or 8).
template <bool INV, int NZ>
__device__ __forceinline__ void dft8(const float2* a, float2* X) {
  constexpr float h = 0.70710678118654752440f;
  float2 e_in[4] = {a[0], a[2], a[4], a[6]};
  float2 o_in[4] = {a[1], a[3], a[5], a[7]};
  float2 E[4], O[4];
  dft4<INV, NZ / 2>(e_in, E);
  dft4<INV, NZ / 2>(o_in, O);
  // W8^1 O1, W8^2 O2, W8^3 O3 (forward W8 = e^{-i pi/4}; inverse conjugate)
  const float2 o1 = INV ? make_float2(h * (O[1].x - O[1].y), h * (O[1].x + O[1].y))
                        : make_float2(h * (O[1].x + O[1].y), h * (O[1].y - O[1].x));
  const float2 o2 = mul_mi<INV>(O[2]);
  const float2 o3 = INV ? make_float2(-h * (O[3].x + O[3].y), h * (O[3].x - O[3].y))
                        : make_float2(h * (O[3].y - O[3].x), -h * (O[3].x + O[3].y));
  X[0] = cadd(E[0], O[0]);
  X[4] = csub(E[0], O[0]);
  X[1] = cadd(E[1], o1);
  X[5] = csub(E[1], o1);
  X[2] = cadd(E[2], o2);
  X[6] = csub(E[2], o2);
  X[3] = cadd(E[3], o3);
  X[7] = csub(E[3], o3);
}

// 3-point DFT: X1, X2 = a0 - (a1 + a2)/2 -/+ i (sqrt(3)/2)(a1 - a2) (forward;
// inverse conjugate).
template <bool INV>
__device__ __forceinline__ void dft3(float2 a0, float2 a1, float2 a2, float2& X0, float2& X1, float2& X2) {
  constexpr float s3 = 0.86602540378443864676f;
  const float2 t = cadd(a1, a2);
  const float2 u = csub(a1, a2);
  X0 = cadd(a0, t);
  const float2 m = make_float2(a0.x - 0.5f * t.x, a0.y - 0.5f * t.y);
  const float2 n = mul_mi<INV>(make_float2(s3 * u.x, s3 * u.y));
  X1 = cadd(m, n);
  X2 = csub(m, n);
}

// 12-point DFT as 3 x 4 (Cooley-Tukey): with j = j1 + 4 j2 and k = ka + 3 kb,
// X[k] = sum_j1 W4^(j1 kb) W12^(j1 ka) sum_j2 W3^(j2 ka) a[j1 + 4 j2]: four
// 3-point DFTs, six constant twiddles, three 4-point DFTs.
template <bool INV>
__device__ __forceinline__ void dft12(const float2* a, float2* X) {
  constexpr float h = 0.5f, s3 = 0.86602540378443864676f;
  // forward W12^e for e = 1, 2, 3, 4, 6 (inverse: their conjugates)
  const float2 w1 = make_float2(s3, -h), w2 = make_float2(h, -s3), w3 = make_float2(0.f, -1.f),
               w4 = make_float2(-h, -s3), w6 = make_float2(-1.f, 0.f);
  float2 A[4][3];
#pragma unroll
  for (int j1 = 0; j1 < 4; ++j1) dft3<INV>(a[j1], a[j1 + 4], a[j1 + 8], A[j1][0], A[j1][1], A[j1][2]);
  A[1][1] = twiddle<INV>(A[1][1], w1);
  A[1][2] = twiddle<INV>(A[1][2], w2);
  A[2][1] = twiddle<INV>(A[2][1], w2);
  A[2][2] = twiddle<INV>(A[2][2], w4);
  A[3][1] = twiddle<INV>(A[3][1], w3);
  A[3][2] = twiddle<INV>(A[3][2], w6);
#pragma unroll
  for (int ka = 0; ka < 3; ++ka) {
    const float2 in[4] = {A[0][ka], A[1][ka], A[2][ka], A[3][ka]};
    float2 out[4];
    dft4<INV, 4>(in, out);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) X[ka + 3 * kb] = out[kb];
  }
}

// Orders one wave's LDS accesses around this point for data that only this
// wave writes and reads (its own slice): a wave's LDS instructions execute in
// issue order, so keeping the compiler from moving accesses across the point
// is enough; the other waves of the block run on without waiting.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-wave LDS slice (float2 elements): pass-A layout R rows of kRowA, pass-B/C
// layout (R*8 tasks) x 9.  Both are padded against bank conflicts: pass B reads
// element b0 + 8 b1 of row k1 = lane >> 3 (b0 = lane & 7), and a row stride of
// 8 (mod 32) float2 puts the eight rows a 16-lane group touches in distinct
// banks (stride 65, 1 mod 32, made two lanes of each group share a bank);
// pass C reads task * 9 + b0, consecutive tasks 9 apart.
constexpr int kRowA = 72;
template <int R>
__host__ __device__ constexpr int wave_lds_elems() {
  return R * kRowA > R * 8 * 9 ? R * kRowA : R * 8 * 9;
}

// Twiddles of one lane for wave_fft, loaded before the data so their latency
// overlaps the data loads instead of stalling the passes: pass A W_K^{lane*k1}
// (k1 < R) and pass B W_K^{R*b0*k2a} (b0 = lane & 7, k2a < 8).  R = 12 has two
// pass-B tasks per lane and reads its twiddles in place.
template <int R>
struct WaveTwiddles {
  float2 a[R];
  float2 b[8];
};

template <int R>
__device__ __forceinline__ WaveTwiddles<R> load_twiddles(const float2* __restrict__ tw, int lane) {
  WaveTwiddles<R> t;
  if constexpr (R <= 8) {
#pragma unroll
    for (int k1 = 1; k1 < R; ++k1) t.a[k1] = tw[lane * k1];
    const int b0 = lane & 7;
#pragma unroll
    for (int k2a = 1; k2a < 8; ++k2a) t.b[k2a] = tw[R * b0 * k2a];
  }
  return t;
}

// Length K = 64*R DFT of one wave, R in {1, 2, 4, 8, 12}.  x[j] = element
// lane + 64 j (elements >= K/2 are zero when HALF_IN; the caller loads zeros
// there, and for even R the radix-R pass skips them).  Output: out[s][k2b] = X[k1 + R*(k2a + 8*k2b)]
// for task = lane + 64 s < 8R, k1 = task >> 3, k2a = task & 7 (only k2b < 4,
// i.e. k < K/2, when HALF_OUT).  tw = W_K^k table.  buf is the wave's own
// slice: the passes synchronize the wave only, so the caller orders the
// slice against other waves' use of the same LDS (barriers before and after).
template <int R, bool INV, bool HALF_IN, bool HALF_OUT>
__device__ __forceinline__ void wave_fft(float2 (&x)[R], float2* buf, const float2* __restrict__ tw,
                                         const WaveTwiddles<R>& twp, int lane,
                                         float2 (&out)[(R * 8 + 63) / 64][8]) {
  constexpr int K = 64 * R;
  constexpr int TPL = (R * 8 + 63) / 64;
  static_assert(R == 1 || R == 2 || R == 4 || R == 8 || R == 12, "unsupported radix");
  // pass A: radix-R over j, twiddle W_K^{lane*k1}
  float2 y[R];
  if constexpr (R == 1) {
    y[0] = x[0];
  } else if constexpr (R == 2) {
    dft2<INV, HALF_IN ? 1 : 2>(x, y);
  } else if constexpr (R == 8) {
    dft8<INV, HALF_IN ? 4 : 8>(x, y);
  } else if constexpr (R == 4) {
    dft4<INV, HALF_IN ? 2 : 4>(x, y);
  } else {
    // R = 12; with HALF_IN the caller's x[6 .. 11] are zero constants
    dft12<INV>(x, y);
  }
#pragma unroll
  for (int k1 = 1; k1 < R; ++k1) y[k1] = twiddle<INV>(y[k1], R <= 8 ? twp.a[k1] : tw[lane * k1]);
#pragma unroll
  for (int k1 = 0; k1 < R; ++k1) buf[k1 * kRowA + lane] = y[k1];
  wave_lds_sync();
  // pass B: task (k1, b0): radix-8 over b1, twiddle W_64^{b0*k2a}
  float2 u[TPL][8];
#pragma unroll
  for (int s = 0; s < TPL; ++s) {
    const int task = lane + 64 * s;
    if (task < R * 8) {
      const int k1 = task >> 3, b0 = task & 7;
#pragma unroll
      for (int b1 = 0; b1 < 8; ++b1) u[s][b1] = buf[k1 * kRowA + b0 + 8 * b1];
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int s = 0; s < TPL; ++s) {
    const int task = lane + 64 * s;
    if (task < R * 8) {
      const int k1 = task >> 3, b0 = task & 7;
      float2 v[8];
      dft8<INV, 8>(u[s], v);
#pragma unroll
      for (int k2a = 1; k2a < 8; ++k2a) v[k2a] = twiddle<INV>(v[k2a], R <= 8 ? twp.b[k2a] : tw[R * b0 * k2a]);
#pragma unroll
      for (int k2a = 0; k2a < 8; ++k2a) buf[(k1 * 8 + k2a) * 9 + b0] = v[k2a];
    }
  }
  wave_lds_sync();
  // pass C: task (k1, k2a): radix-8 over b0
#pragma unroll
  for (int s = 0; s < TPL; ++s) {
    const int task = lane + 64 * s;
    if (task < R * 8) {
      float2 v[8];
#pragma unroll
      for (int b0 = 0; b0 < 8; ++b0) v[b0] = buf[task * 9 + b0];
      dft8<INV, 8>(v, out[s]);
    }
  }
  (void)K;
  (void)HALF_OUT;
}

// Block exchange area: row k holds one element of each of the block's 8
// waves (coils).  Element c of row k sits in slot c ^ ((k ^ k >> 3) & 7), so
// the three access patterns (a wave walking rows, a thread walking the 8
// elements of its row, contiguous fills from global memory) stay at the
// minimum 2-way bank sharing instead of up to 16-way conflicts.
__device__ __forceinline__ int xslot(int k, int c) { return k * 8 + (c ^ ((k ^ (k >> 3)) & 7)); }

// ROWS rows x 8 coils of global memory held in registers on their way to the
// exchange area (512-thread blocks): thread t holds elements t + 512 i, row
// e >> 3, coil e & 7.
template <int ROWS>
struct ExchangeRegs {
  float2 v[(ROWS * 8 + 511) / 512];
};

// Issue the loads of ROWS rows (row r, coil c at src[r * row_stride + c];
// coils >= cw read as zero) without waiting for them.
template <int ROWS>
__device__ __forceinline__ void load_exchange(const float2* __restrict__ src, uint64_t row_stride, uint32_t cw,
                                              ExchangeRegs<ROWS>& r) {
  constexpr int kElems = ROWS * 8;
#pragma unroll
  for (int i = 0; i < (kElems + 511) / 512; ++i) {
    const uint32_t e = threadIdx.x + 512u * i;
    const uint32_t row = e >> 3, c = e & 7;
    r.v[i] = (e < static_cast<uint32_t>(kElems) && c < cw) ? src[row * row_stride + c] : make_float2(0.f, 0.f);
  }
}

// Write loaded rows to the exchange area.
template <int ROWS>
__device__ __forceinline__ void write_exchange(const ExchangeRegs<ROWS>& r, float2* lds) {
  constexpr int kElems = ROWS * 8;
#pragma unroll
  for (int i = 0; i < (kElems + 511) / 512; ++i) {
    const uint32_t e = threadIdx.x + 512u * i;
    if (e < static_cast<uint32_t>(kElems)) lds[xslot(static_cast<int>(e >> 3), static_cast<int>(e & 7))] = r.v[i];
  }
}

// Fill the exchange area with ROWS rows x 8 coils from global memory: all
// loads are issued before the first LDS write, so a block pays one memory
// latency, not one per loop trip.  512-thread blocks.
template <int ROWS>
__device__ __forceinline__ void fill_exchange(const float2* __restrict__ src, uint64_t row_stride, uint32_t cw,
                                              float2* lds) {
  ExchangeRegs<ROWS> r;
  load_exchange<ROWS>(src, row_stride, cw, r);
  write_exchange<ROWS>(r, lds);
}

// Store the first `slots` coils of ROWS rows of the exchange area to dst + r
// * row_stride (float2 units; slots <= 8; exchange slots past a chunk's
// coils hold zeros).  An even slot count at an even stride (16-byte aligned
// rows) goes out as float4 coil pairs, four neighbouring threads per row, so
// a wave's store covers 16 whole rows; otherwise one thread writes a row coil
// by coil.  512-thread blocks.
template <int ROWS>
__device__ __forceinline__ void store_rows(float2* __restrict__ dst, uint64_t row_stride, uint32_t slots,
                                           const float2* lds) {
  if ((slots & 1u) == 0 && (row_stride & 1u) == 0) {
#pragma unroll
    for (uint32_t e = threadIdx.x; e < static_cast<uint32_t>(ROWS) * 4; e += 512u) {
      const uint32_t r = e >> 2, q = e & 3;
      if (2 * q >= slots) continue;
      const float2 u = lds[xslot(static_cast<int>(r), static_cast<int>(2 * q))];
      const float2 v = lds[xslot(static_cast<int>(r), static_cast<int>(2 * q + 1))];
      reinterpret_cast<float4*>(dst + r * row_stride)[q] = make_float4(u.x, u.y, v.x, v.y);
    }
  } else {
    for (uint32_t r = threadIdx.x; r < static_cast<uint32_t>(ROWS); r += 512u)
      for (uint32_t i = 0; i < slots; ++i)
        dst[r * row_stride + i] = lds[xslot(static_cast<int>(r), static_cast<int>(i))];
  }
}

// Write the wave's outputs to the block exchange area, element w of row k
// (k < K, or k < K/2 when HALF_OUT).  Call between two __syncthreads.
template <int R, bool HALF_OUT>
__device__ __forceinline__ void emit_exchange(const float2 (&out)[(R * 8 + 63) / 64][8], float2* lds,
                                              int lane, int w) {
  constexpr int TPL = (R * 8 + 63) / 64;
#pragma unroll
  for (int s = 0; s < TPL; ++s) {
    const int task = lane + 64 * s;
    if (task >= R * 8) continue;
    const int k1 = task >> 3, k2a = task & 7;
#pragma unroll
    for (int k2b = 0; k2b < (HALF_OUT ? 4 : 8); ++k2b) {
      const int k = k1 + R * (k2a + 8 * k2b);
      lds[xslot(k, w)] = out[s][k2b];
    }
  }
}

// problems > 1: batched launch over blockIdx.z (strides and coils_of in a).
int fft_forward(const FftArgs& a, int out_kind, hipStream_t s, uint32_t problems = 1);
int fft_inverse(const FftArgs& a, hipStream_t s, uint32_t problems = 1);
bool fft_length_supported(uint32_t k);

}  // namespace tsn
