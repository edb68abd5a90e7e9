// mlp.hip -- the learned surrogate of the spectral dyadic: a fully connected
// network (SiLU hidden layers) trained and evaluated on the gfx950 matrix
// cores in BF16 with FP32 accumulation.  No reference counterpart (the
// reference computes every node with the exact solver, which here stays the
// fallback and the check on the surrogate).
//
// Every layer product is one kernel, k_gemm: C(m, n) = sum_k A(m, k) B(n, k)
// on v_mfma_f32_32x32x16_bf16, with either operand stored k-fastest or
// transposed, and a fused epilogue:
//   forward   Z = H W^T + b  (A = H [batch][k], B = W [n][k]); store Z (FP32)
//             and H' = bf16(SiLU(Z)), or Y = Z for the output layer;
//   backward  dH = dZ W      (A = dZ [batch][n], B = W read transposed), times
//             SiLU'(Z) -> bf16 dZ of the layer below;
//             dW = dZ^T H    (A = dZ read transposed, B = H read transposed,
//             the batch is the reduction index) in FP32.
// Shapes are padded to 64 (rows, columns) and 32 (reduction) with zeros.
#include <hip/hip_runtime.h>

#include "ytf_kernels.h"

namespace ytf {

typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int TM = 64, TN = 64, TK = 32, LK = TK + 8;

// SiLU and its derivative from the hardware exp2 and reciprocal (1 ulp each;
// the results are rounded to bf16): sigma(z) = 1 / (1 + 2^(-z log2 e)).
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

template <int EPI, bool AT, bool BT>
__global__ __launch_bounds__(256) void k_gemm(MlpGemm p) {
  __shared__ __bf16 sA[TM * LK], sB[TN * LK];
  const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
  // split-K: slice z of the reduction writes its own partial sum
  const int kz = blockIdx.z, kslice = p.K / gridDim.z;
  const int kbeg = kz * kslice, kend = kbeg + kslice;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 31, lk = (lane >> 5) * 8;
  f16v acc;
#pragma unroll
  for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    // stage 64 x 32 of each operand: 256 chunks of 8, one per thread
    if (!AT) {
      const int row = tid >> 2, kc = (tid & 3) * 8;
      *(b8 *)(sA + row * LK + kc) = *(const b8 *)(p.A + (int64_t)(m0 + row) * p.lda + k0 + kc);
    } else {
      const int kk = tid >> 3, mc = (tid & 7) * 8;
      const b8 v = *(const b8 *)(p.A + (int64_t)(k0 + kk) * p.lda + m0 + mc);
#pragma unroll
      for (int j = 0; j < 8; ++j) sA[(mc + j) * LK + kk] = v[j];
    }
    if (!BT) {
      const int row = tid >> 2, kc = (tid & 3) * 8;
      *(b8 *)(sB + row * LK + kc) = *(const b8 *)(p.B + (int64_t)(n0 + row) * p.ldb + k0 + kc);
    } else {
      const int kk = tid >> 3, nc = (tid & 7) * 8;
      const b8 v = *(const b8 *)(p.B + (int64_t)(k0 + kk) * p.ldb + n0 + nc);
#pragma unroll
      for (int j = 0; j < 8; ++j) sB[(nc + j) * LK + kk] = v[j];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < TK / 16; ++s) {
      const b8 a = *(const b8 *)(sA + (wm * 32 + lr) * LK + s * 16 + lk);
      const b8 b = *(const b8 *)(sB + (wn * 32 + lr) * LK + s * 16 + lk);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // D layout: lane holds column lane % 32, rows 8 g + 4 (lane / 32) + e
  const int col = n0 + wn * 32 + lr;
  const int hi = lane >> 5;
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = m0 + wm * 32 + 8 * g + 4 * hi + e;
      const int64_t o = (int64_t)row * p.ldc + col;
      const float v = acc[4 * g + e];
      if (EPI == MLP_EPI_F32) {
        p.C32[o + (int64_t)kz * p.split_stride] = v;
      } else if (EPI == MLP_EPI_BIAS_F32) {
        p.C32[o] = v + p.bias[col];
      } else if (EPI == MLP_EPI_BIAS_SILU) {
        const float z = v + p.bias[col];
        if (p.C32) p.C32[o] = z;
        p.C16[o] = (__bf16)silu(z);
      } else {  // MLP_EPI_DSILU
        p.C16[o] = (__bf16)(v * dsilu(p.Zin[o]));
      }
    }
}

hipError_t launch_mlp_gemm(const MlpGemm &p, int epi, bool at, bool bt, hipStream_t s) {
  if (p.M % TM || p.N % TN || p.K % TK || p.M <= 0 || p.N <= 0 || p.K <= 0) return hipErrorInvalidValue;
  const int splits = p.ksplit > 1 ? p.ksplit : 1;
  if (splits > 1 && (epi != MLP_EPI_F32 || p.K % (TK * splits))) return hipErrorInvalidValue;
  dim3 grid(p.M / TM, p.N / TN, splits);
#define YTF_G(E, X, Y) hipLaunchKernelGGL((k_gemm<E, X, Y>), grid, dim3(256), 0, s, p)
  if (epi == MLP_EPI_BIAS_SILU && !at && !bt) YTF_G(MLP_EPI_BIAS_SILU, false, false);
  else if (epi == MLP_EPI_BIAS_F32 && !at && !bt) YTF_G(MLP_EPI_BIAS_F32, false, false);
  else if (epi == MLP_EPI_DSILU && !at && bt) YTF_G(MLP_EPI_DSILU, false, true);
  else if (epi == MLP_EPI_F32 && !at && bt) YTF_G(MLP_EPI_F32, false, true);
  else if (epi == MLP_EPI_F32 && at && bt) YTF_G(MLP_EPI_F32, true, true);
  else return hipErrorInvalidValue;
#undef YTF_G
  return hipGetLastError();
}

// dst[i] = sum over z of part[z * stride + i] (fixed order)
__global__ void k_mlp_reduce(int64_t n, int splits, int64_t stride, const float *__restrict__ part,
                             float *__restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int z = 0; z < splits; ++z) s += part[z * stride + i];
  dst[i] = s;
}

// raw inputs x[n][nin] -> bf16 X[rows][ldx] = (x - shift) * scale, zero
// padded.  With kz features, two more columns follow the inputs: with
// u^2 = x0^2 + x1^2 (the in-plane wavevector over omega), Re and Im of
// sqrt(1 - u^2), i.e. kz/omega of the air on its principal branch, which makes
// the reflection coefficients smooth across the light line u = 1.
__global__ void k_mlp_inputs(int64_t n, int64_t rows, int nin, int ldx, int kzf, const float *__restrict__ x,
                             const float *__restrict__ shift, const float *__restrict__ scale, __bf16 *__restrict__ X) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= rows * ldx) return;
  const int64_t r = t / ldx;
  const int c = (int)(t % ldx);
  float v = 0.0f;
  if (r < n && c < nin) {
    v = (x[r * nin + c] - shift[c]) * scale[c];
  } else if (r < n && kzf && c < nin + 2) {
    const float u2 = x[r * nin] * x[r * nin] + x[r * nin + 1] * x[r * nin + 1];
    v = c == nin ? sqrtf(fmaxf(1.0f - u2, 0.0f)) : sqrtf(fmaxf(u2 - 1.0f, 0.0f));
  }
  X[t] = (__bf16)v;
}

// MSE loss of the normalised outputs and its gradient:
//   t = (y - out_shift) / out_scale,  L = sum (Y - t)^2 / (n nout),
//   dY = 2 (Y - t) / (n nout) -> bf16 dZ (zero in padding).
__global__ void k_mlp_mse(int64_t n, int64_t rows, int nout, int ldy, const float *__restrict__ Y,
                          const float *__restrict__ y, const float *__restrict__ oshift,
                          const float *__restrict__ oscale, __bf16 *__restrict__ dZ, float *__restrict__ loss) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  float sq = 0.0f;
  if (t < rows * ldy) {
    const int64_t r = t / ldy;
    const int c = (int)(t % ldy);
    float d = 0.0f;
    if (r < n && c < nout) {
      const float tv = (y[r * nout + c] - oshift[c]) / oscale[c];
      d = Y[t] - tv;
      sq = d * d;
    }
    dZ[t] = (__bf16)(2.0f * d / (float)(n * nout));
  }
  if (loss) {
    __shared__ float red[256];
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)(n * nout));
  }
}

// Column sums of dZ [rows][ld] in two deterministic stages: block (x, z)
// sums 64 columns over row slice z (4 row groups per block, coalesced 64-column
// rows), part[z][c]; then k_mlp_reduce adds the slices in order.
__global__ __launch_bounds__(256) void k_mlp_colsum(int64_t rows, int ld, int slices, const __bf16 *__restrict__ dZ,
                                                    float *__restrict__ part) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
  const int64_t per = (rows + slices - 1) / slices;
  const int64_t r0 = blockIdx.y * per, r1 = r0 + per < rows ? r0 + per : rows;
  float s = 0.0f;
  if (c < ld)
    for (int64_t r = r0 + g; r < r1; r += 4) s += (float)dZ[r * ld + c];
  red[g][threadIdx.x & 63] = s;
  __syncthreads();
  if (g == 0 && c < ld) part[(int64_t)blockIdx.y * ld + c] = red[0][threadIdx.x] + red[1][threadIdx.x] +
                                                              red[2][threadIdx.x] + red[3][threadIdx.x];
}

// Y[rows][ldy] (normalised) -> y[n][nout] raw
__global__ void k_mlp_outputs(int64_t n, int nout, int ldy, const float *__restrict__ Y,
                              const float *__restrict__ oshift, const float *__restrict__ oscale,
                              float *__restrict__ y) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * nout) return;
  const int64_t r = t / nout;
  const int c = (int)(t % nout);
  y[t] = Y[r * ldy + c] * oscale[c] + oshift[c];
}

// Adam on the FP32 master parameters; refreshes the bf16 copy.
__global__ void k_mlp_adam(int64_t n, float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ g, __bf16 *__restrict__ w16, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (w16) w16[i] = (__bf16)wi;
}

__global__ void k_mlp_to_bf16(int64_t n, const float *__restrict__ w, __bf16 *__restrict__ w16) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) w16[i] = (__bf16)w[i];
}

__global__ void k_mlp_zero(int64_t n, float *__restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0f;
}

static unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_mlp_inputs(int64_t n, int64_t rows, int nin, int ldx, int kzf, const float *x, const float *shift,
                             const float *scale, void *X, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_inputs, dim3(blocks_for(rows * ldx)), dim3(256), 0, s, n, rows, nin, ldx, kzf, x, shift,
                     scale, (__bf16 *)X);
  return hipGetLastError();
}
hipError_t launch_mlp_mse(int64_t n, int64_t rows, int nout, int ldy, const float *Y, const float *y,
                          const float *oshift, const float *oscale, void *dZ, float *loss, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_mse, dim3(blocks_for(rows * ldy)), dim3(256), 0, s, n, rows, nout, ldy, Y, y, oshift,
                     oscale, (__bf16 *)dZ, loss);
  return hipGetLastError();
}
hipError_t launch_mlp_colsum(int64_t rows, int ld, const void *dZ, float *part, float *db, hipStream_t s) {
  int slices = (int)((rows + 255) / 256);
  if (slices > 64) slices = 64;
  if (slices < 1) slices = 1;
  hipLaunchKernelGGL(k_mlp_colsum, dim3((ld + 63) / 64, slices), dim3(256), 0, s, rows, ld, slices,
                     (const __bf16 *)dZ, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((ld + 255) / 256), dim3(256), 0, s, (int64_t)ld, slices, (int64_t)ld,
                     (const float *)part, db);
  return hipGetLastError();
}
hipError_t launch_mlp_outputs(int64_t n, int nout, int ldy, const float *Y, const float *oshift, const float *oscale,
                              float *y, hipStream_t s) {
  if (n * nout == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_outputs, dim3(blocks_for(n * nout)), dim3(256), 0, s, n, nout, ldy, Y, oshift, oscale, y);
  return hipGetLastError();
}
hipError_t launch_mlp_adam(int64_t n, float *w, float *m, float *v, const float *g, void *w16, float lr, float b1,
                           float b2, float eps, float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(n)), dim3(256), 0, s, n, w, m, v, g, (__bf16 *)w16, lr, b1, b2, eps,
                     c1, c2);
  return hipGetLastError();
}
hipError_t launch_mlp_to_bf16(int64_t n, const float *w, void *w16, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_to_bf16, dim3(blocks_for(n)), dim3(256), 0, s, n, w, (__bf16 *)w16);
  return hipGetLastError();
}
hipError_t launch_mlp_reduce(int64_t n, int splits, int64_t stride, const float *part, float *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks_for(n)), dim3(256), 0, s, n, splits, stride, part, dst);
  return hipGetLastError();
}
hipError_t launch_mlp_zero(int64_t n, float *x, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_zero, dim3(blocks_for(n)), dim3(256), 0, s, n, x);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Whole-network forward in one kernel.  A block takes FB = 128 rows; their
// activations stay in LDS as bf16 [row][feature] (row stride WIDTH + 8).
// Every layer is computed transposed, C^T = W H^T, on
// v_mfma_f32_32x32x16_bf16 with the weights (A operand) read from global
// memory in fragment order (launch_mlp_pack: one coalesced 1 KiB load per wave,
// 32-feature tile and 16-wide k step, two steps ahead; L2-resident, since
// every block reads the same weights) and the activations (B operand) from LDS: the accumulator of a lane then holds 4
// consecutive features of one row, stored back as one 8-byte LDS write.
// NW waves per block (8 for width 256: 4 waves per SIMD at two blocks per CU,
// which hides the LDS and weight-load latency better than 4 waves per block;
// 4 for width 128).  Hidden layer: wave w owns features [w W/NW, (w + 1) W/NW)
// (W/(32 NW) tiles of 32) for all 4 row tiles; output layer: the 4 ceil(nout / 32)
// (feature tile, row tile) pairs spread over the waves.  The sums run over k in 16-wide steps in increasing order,
// as in k_gemm, and the epilogues (bias, SiLU, output scaling) are the same.
// Two variants:
//   generic:   x[n][nin] (normalised in-kernel, light-line features) -> y[n][nout];
//   surrogate: inputs (kx/omega, ky/omega, constants); the 18 outputs t go
//              through LDS to one thread per row, which forms
//              g~_s = t omega (1 + k^2/omega^2) exp(i kz0 zsum) / kz0 in FP64
//              ([9][n] complex, entry c + 3q).  The rows flagged from their
//              inputs alone (|k - omega| < branch_tol omega or k < origin_tol
//              omega) are already in the flag array (k_sur_preflag, launched
//              before; with skip_flagged their g~ is left to the exact solve);
//              a row not among them with |t|_2 > tbig is flagged here and
//              appended to the flagged-row list.
constexpr int FB = 128;
#ifndef MLP_NW256
#define MLP_NW256 8  // waves per block of the width-256 network
#endif

typedef __bf16 b4 __attribute__((ext_vector_type(4)));

// Epilogue of one 4-feature group of a hidden layer: h = bf16(SiLU(z + b)) for
// features f .. f + 3 of LDS row `row` (one 8-byte store).
template <int LA>
__device__ __forceinline__ void epi4(__bf16 *act, int row, int f, const f16v &z, int g, const float *b) {
  b4 h;
#pragma unroll
  for (int e = 0; e < 4; ++e) h[e] = (__bf16)silu(z[4 * g + e] + b[f + e]);
  *(b4 *)(act + row * LA + f) = h;
}

// One half-layer phase of the fused forward.  The wave's accumulators accM
// (FT feature tiles x the 2 row tiles of half `hm`) take the products of the
// layer whose packed weights start at Wp (K = WIDTH, WIDTH / 16 steps), while
// the epilogue of the other half's finished accumulators accE (bias bE, rows
// of half `he`) is spread over the steps, one 4-feature group per step, so
// the SiLU arithmetic overlaps the matrix-core work.  do_epi = false: products
// only.
template <int WIDTH, int FT, int LA>
__device__ __forceinline__ void half_phase(const b8 *Wp, const __bf16 *actr, f16v (&accM)[FT][2], int hm,
                                           f16v (&accE)[FT][2], int he, const float *bE, bool do_epi, __bf16 *act,
                                           int f0, int lr, int lk, int hi) {
  constexpr int NKS = WIDTH / 16;
  constexpr int NG = FT * 2 * 4;  // epilogue groups of the other half
  static_assert(NKS % NG == 0, "epilogue groups spread evenly over the k steps");
#pragma unroll
  for (int ft = 0; ft < FT; ++ft)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) accM[ft][j][e] = 0.0f;
  b8 a0[FT], a1[FT];
#pragma unroll
  for (int ft = 0; ft < FT; ++ft) {
    a0[ft] = Wp[ft * NKS * 64];
    a1[ft] = Wp[(ft * NKS + 1) * 64];
  }
#pragma unroll
  for (int k = 0; k < NKS; ++k) {
    b8 a2[FT], bb[2];
#pragma unroll
    for (int ft = 0; ft < FT; ++ft) a2[ft] = k + 2 < NKS ? Wp[(ft * NKS + k + 2) * 64] : a1[ft];
#pragma unroll
    for (int j = 0; j < 2; ++j) bb[j] = *(const b8 *)(actr + ((2 * hm + j) * 32 + lr) * LA + 16 * k + lk);
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        accM[ft][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0[ft], bb[j], accM[ft][j], 0, 0, 0);
    if (do_epi && k % (NKS / NG) == 0) {  // group q of the other half: tile (q / 8, (q / 4) % 2), g = q % 4
      const int q = k / (NKS / NG);
      const int ft = q / 8, j = (q / 4) % 2, g = q % 4;
      epi4<LA>(act, (2 * he + j) * 32 + lr, f0 + ft * 32 + 8 * g + 4 * hi, accE[ft][j], g, bE);
    }
#pragma unroll
    for (int ft = 0; ft < FT; ++ft) {
      a0[ft] = a1[ft];
      a1[ft] = a2[ft];
    }
  }
}

// The full epilogue of one half (no overlapping products).
template <int FT, int LA>
__device__ __forceinline__ void half_epilogue(__bf16 *act, f16v (&acc)[FT][2], int h, const float *b, int f0, int lr,
                                              int hi) {
#pragma unroll
  for (int ft = 0; ft < FT; ++ft)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int g = 0; g < 4; ++g) epi4<LA>(act, (2 * h + j) * 32 + lr, f0 + ft * 32 + 8 * g + 4 * hi, acc[ft][j], g, b);
}

template <int WIDTH, bool SUR, int NW>
__global__ __launch_bounds__(64 * NW, 2) void k_mlp_fused(MlpFused p) {
  constexpr int NT = 64 * NW;         // threads per block
  constexpr int LA = WIDTH + 8;
  constexpr int FT = WIDTH / (32 * NW);  // 32-feature tiles per wave in a hidden layer
  static_assert(FT >= 1, "at least one feature tile per wave");
  constexpr int TS = 20;  // floats per row of the surrogate's output staging
  static_assert(FB * TS * 4 <= FB * LA * 2, "output staging fits the activations");
  __shared__ __bf16 act[FB * LA];
  __shared__ float bsh[(MLP_FUSED_MAX_LAYERS - 1) * WIDTH];  // hidden-layer biases (read without vmcnt waits)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 31, hi = lane >> 5, lk = hi * 8;
  const int64_t r0 = (int64_t)blockIdx.x * FB;
  const float *ish = p.norm, *isc = p.norm + p.nin;
  const float *osh = p.norm + 2 * p.nin, *osc = osh + p.nout;
  const int kin = p.nin + 2 * p.kzf;
  const int k0e = (kin + 15) / 16 * 16;

  for (int e = tid; e < (p.nl - 1) * WIDTH; e += NT) bsh[e] = p.P32[p.offb[e / WIDTH] + e % WIDTH];
  // layer-0 inputs, zero padded to k0e columns
  for (int e = tid; e < FB * k0e; e += NT) {
    const int r = e / k0e, c = e % k0e;
    const int64_t row = r0 + r;
    float v = 0.0f;
    if (row < p.n && c < kin) {
      float x0, x1;
      if (SUR) {
        x0 = (float)(p.kx[row] / p.omega);
        x1 = (float)(p.ky[row] / p.omega);
      } else {
        x0 = p.x[row * p.nin];
        x1 = p.nin > 1 ? p.x[row * p.nin + 1] : 0.0f;
      }
      if (c < p.nin) {
        const float xc = c == 0 ? x0 : (c == 1 ? x1 : (SUR ? p.xconst[c - 2] : p.x[row * p.nin + c]));
        v = (xc - ish[c]) * isc[c];
      } else {
        const float u2 = x0 * x0 + x1 * x1;
        v = c == p.nin ? sqrtf(fmaxf(1.0f - u2, 0.0f)) : sqrtf(fmaxf(u2 - 1.0f, 0.0f));
      }
    }
    act[r * LA + c] = (__bf16)v;
  }
  __syncthreads();

  // hidden layers, in two halves of 64 rows (A: row tiles 0-1, B: 2-3) so
  // that the epilogue of one half overlaps the products of the other:
  //   layer 0: products of A and B; epilogue A;
  //   layer l >= 1: [products l of A | epilogue l-1 of B], [products l of B | epilogue l of A];
  //   then the epilogue of B of the last hidden layer.
  // Every bracket ends with a barrier: a half's rows are rewritten only after
  // every wave has read them, and read only after every wave has written them.
  const int f0 = wave * (WIDTH / NW);
  f16v accA[FT][2], accB[FT][2];
  {
    const int ksl = p.kp0 / 16, nks = k0e / 16;
    const b8 *Wp = (const b8 *)(p.W16 + p.offW[0]) + (int64_t)(f0 / 32) * ksl * 64 + lane;
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) accA[ft][j][e] = accB[ft][j][e] = 0.0f;
    for (int k = 0; k < nks; ++k) {
      b8 a[FT], bb[4];
#pragma unroll
      for (int ft = 0; ft < FT; ++ft) a[ft] = Wp[((int64_t)ft * ksl + k) * 64];
#pragma unroll
      for (int st = 0; st < 4; ++st) bb[st] = *(const b8 *)(act + (st * 32 + lr) * LA + 16 * k + lk);
#pragma unroll
      for (int ft = 0; ft < FT; ++ft)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accA[ft][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ft], bb[j], accA[ft][j], 0, 0, 0);
          accB[ft][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ft], bb[2 + j], accB[ft][j], 0, 0, 0);
        }
    }
  }
  __syncthreads();
  half_epilogue<FT, LA>(act, accA, 0, bsh, f0, lr, hi);
  __syncthreads();
  for (int l = 1; l < p.nl - 1; ++l) {
    const b8 *Wp = (const b8 *)(p.W16 + p.offW[l]) + (int64_t)(f0 / 32) * (WIDTH / 16) * 64 + lane;
    half_phase<WIDTH, FT, LA>(Wp, act, accA, 0, accB, 1, bsh + (l - 1) * WIDTH, true, act, f0, lr, lk, hi);
    __syncthreads();
    half_phase<WIDTH, FT, LA>(Wp, act, accB, 1, accA, 0, bsh + l * WIDTH, true, act, f0, lr, lk, hi);
    __syncthreads();
  }
  half_epilogue<FT, LA>(act, accB, 1, bsh + (p.nl - 2) * WIDTH, f0, lr, hi);
  __syncthreads();

  // output layer
  {
    const int l = p.nl - 1;
    const __bf16 *W = p.W16 + p.offW[l];
    const float *b = p.P32 + p.offb[l];
    // tiles t = 4 ft + st (nt feature tiles x 4 row tiles), ceil(4 nt / NW) consecutive ones per wave
    const int nt = (p.nout + 31) / 32;
    const int per = (4 * nt + NW - 1) / NW;
    const int t0 = wave * per;
    const int nst = t0 >= 4 * nt ? 0 : min(per, 4 * nt - t0);
    const int ft = t0 / 4, st0 = t0 % 4;
    f16v acc[2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[j][e] = 0.0f;
    const b8 *Wp = (const b8 *)W + (int64_t)(nst ? ft : 0) * (WIDTH / 16) * 64 + lane;
    b8 a0 = Wp[0], a1 = Wp[64];
    for (int k = 0; k < WIDTH / 16; ++k) {
      const b8 a2 = k + 2 < WIDTH / 16 ? Wp[(k + 2) * 64] : a1;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (j < nst) {
          const b8 bb = *(const b8 *)(act + ((st0 + j) * 32 + lr) * LA + 16 * k + lk);
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bb, acc[j], 0, 0, 0);
        }
      }
      a0 = a1;
      a1 = a2;
    }
    __syncthreads();
    float *tb = (float *)act;  // surrogate: [FB][TS] staged outputs
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (j >= nst) continue;
      const int s = (st0 + j) * 32 + lr;
      const int64_t row = r0 + s;
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int f = ft * 32 + 8 * g + 4 * hi + e;
          if (f < p.nout) {
            const float v = (acc[j][4 * g + e] + b[f]) * osc[f] + osh[f];
            if (SUR) tb[s * TS + f] = v;
            else if (row < p.n) p.y[row * p.nout + f] = v;
          }
        }
    }
    if (SUR) {
      __syncthreads();
      const int64_t row = r0 + tid;
      if (tid < FB && row < p.n) {
        const float *t = tb + tid * TS;
        double tn2 = 0.0;
#pragma unroll
        for (int c = 0; c < 18; ++c) tn2 += (double)t[c] * (double)t[c];
        const double kx = p.kx[row], ky = p.ky[row], w = p.omega;
        const double k2 = kx * kx + ky * ky;
        const bool pre = p.flags[row] != 0;
        const bool big = !pre && sqrt(tn2) > p.tbig;
        const cd kz = csqrt(cd{w * w - k2, 0.0});
        const cd P = cexp(cd{-kz.i * p.zsum, kz.r * p.zsum});
        const cd fac = (cd{w * P.r, w * P.i} * cd{1.0 + k2 / (w * w), 0.0}) / kz;
        if (!(pre && p.skip_flagged)) {
#pragma unroll
          for (int e = 0; e < 9; ++e) {
            const cd v = cd{(double)t[e], (double)t[9 + e]} * fac;
            p.ghat[2 * (e * p.n + row)] = v.r;
            p.ghat[2 * (e * p.n + row) + 1] = v.i;
          }
        }
        if (big) {
          p.flags[row] = 1;
          p.flagged[atomicAdd(p.n_flagged, 1)] = (int)row;
        }
      }
    }
  }
}

__global__ void k_mlp_pack(int N, int K, const __bf16 *__restrict__ W, __bf16 *__restrict__ Wpk) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // one fragment lane (8 values)
  const int64_t nfrag = (int64_t)(N / 32) * (K / 16) * 64;
  if (t >= nfrag) return;
  const int lane = (int)(t % 64);
  const int64_t ts = t / 64;
  const int step = (int)(ts % (K / 16)), tile = (int)(ts / (K / 16));
  const int row = 32 * tile + lane % 32, col = 16 * step + 8 * (lane / 32);
  *(b8 *)(Wpk + 8 * t) = *(const b8 *)(W + (int64_t)row * K + col);
}

hipError_t launch_mlp_pack(int N, int K, const __bf16 *W, __bf16 *Wpk, hipStream_t s) {
  if (N % 32 || K % 16 || N <= 0 || K <= 0) return hipErrorInvalidValue;
  const int64_t nfrag = (int64_t)(N / 32) * (K / 16) * 64;
  hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((nfrag + 255) / 256)), dim3(256), 0, s, N, K, W, Wpk);
  return hipGetLastError();
}

bool mlp_fused_supported(int width, int nin_total, int nout, int nl) {
  return (width == 128 || width == 256) && nin_total <= 64 && nout >= 1 && nout <= 64 && nl >= 2 &&
         nl <= MLP_FUSED_MAX_LAYERS;
}

hipError_t launch_mlp_fused(const MlpFused &p, int width, bool surrogate, hipStream_t s) {
  if (!mlp_fused_supported(width, p.nin + 2 * p.kzf, p.nout, p.nl) || (surrogate && p.nout != 18))
    return hipErrorInvalidValue;
  if (p.n <= 0) return hipSuccess;
  const dim3 grid((unsigned)((p.n + FB - 1) / FB));
  if (width == 256) {
    if (surrogate) hipLaunchKernelGGL((k_mlp_fused<256, true, MLP_NW256>), grid, dim3(64 * MLP_NW256), 0, s, p);
    else hipLaunchKernelGGL((k_mlp_fused<256, false, MLP_NW256>), grid, dim3(64 * MLP_NW256), 0, s, p);
  } else {
    if (surrogate) hipLaunchKernelGGL((k_mlp_fused<128, true, 4>), grid, dim3(256), 0, s, p);
    else hipLaunchKernelGGL((k_mlp_fused<128, false, 4>), grid, dim3(256), 0, s, p);
  }
  return hipGetLastError();
}

// Input flags of the surrogate rows: flags[i] = 1 where |k - omega| <
// branch_tol omega or k < origin_tol omega (near the light line or the
// origin) or k lies in one of the nwin pole windows [lo, hi] (win: lo0 hi0
// lo1 hi1 ...), else 0; the flagged rows are appended to list (one counter
// increment per wave).
__global__ void k_sur_preflag(int64_t n, const double *__restrict__ kx, const double *__restrict__ ky, double w,
                              double branch_tol, double origin_tol, const double *__restrict__ win, int nwin,
                              unsigned char *__restrict__ flags, int *__restrict__ count, int *__restrict__ list) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool f = false;
  if (i < n) {
    const double a = kx[i], b = ky[i];
    const double kr = sqrt(a * a + b * b);
    f = fabs(kr - w) < branch_tol * w || kr < origin_tol * w;
    for (int t = 0; t < nwin; ++t) f = f || (kr >= win[2 * t] && kr <= win[2 * t + 1]);
    flags[i] = f ? 1 : 0;
  }
  const unsigned long long bal = __ballot(f);
  if (bal == 0) return;
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0) base = atomicAdd(count, __popcll(bal));
  base = __shfl(base, 0);
  if (f) list[base + __popcll(bal & ((1ull << lane) - 1))] = (int)i;
}

hipError_t launch_sur_preflag(int64_t n, const double *kx, const double *ky, double w, double branch_tol,
                              double origin_tol, const double *win, int nwin, unsigned char *flags, int *count,
                              int *list, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sur_preflag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, kx, ky, w, branch_tol,
                     origin_tol, win, nwin, flags, count, list);
  return hipGetLastError();
}

}  // namespace ytf
