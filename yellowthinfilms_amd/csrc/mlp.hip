// mlp.hip -- the learned surrogate of the spectral dyadic: a fully connected
// network (SiLU hidden layers) trained and evaluated on the gfx950 matrix
// cores in BF16 with FP32 accumulation.  No reference counterpart (the
// reference computes every node with the exact solver, which here stays the
// fallback and the judge of the surrogate).
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

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
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

}  // namespace ytf
