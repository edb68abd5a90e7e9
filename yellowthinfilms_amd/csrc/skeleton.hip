// skeleton.hip -- z skeletons of the operator's blocks without z structure
// (_zskeleton, mod_GreenOperator.jl:262-302), FP64.
//
// A side (the source planes of the source layer, or the observation planes of
// one layer) has the exponential family of its layer: one row per node term,
//     A[r, j] = exp(i Q_r (sigma_f z_j + c_f)),   r = (entry, node),
// with Q_r the node's mode wavenumber of entry e of a [4][n] array (Qf1 Qf2
// Qb1 Qb2 of the source layer, or KZf1 KZf2 KZb1 KZb2 of the observation
// layer), family f = fam[e] (or skipped), and exp underflowing to 0 below
// exp(-700).  Zero and non-finite Q (invalid nodes) give no row (_qlist).
//
//   k_sk_tsqr: the triangular factor R (N x N) of A by a tall-skinny QR.
//     Every workgroup keeps its R in LDS, stacks chunks of 64 rows under it
//     and re-factors with Householder reflections; a second launch with one
//     workgroup factors the stacked partial R's.  R is the reference's R up
//     to a unitary diagonal row scaling, so the column-pivoted QR of R (on the
//     host) picks the same planes.
//   k_sk_err: max_r,j |A[r, j] - sum_q A[r, idx_q] P[j, q]| of an
//     interpolation (idx, P), for the rank bisection.
#include <hip/hip_runtime.h>

#include "ytf_kernels.h"

namespace ytf {

__device__ __forceinline__ cd sk_exp(cd q, double t) {
  // exp(i q t) with underflow to 0 below exp(-700) (_family_rows)
  const double re = -q.i * t, im = q.r * t;
  if (re < -700.0) return cd{0.0, 0.0};
  return cexp(cd{re, im});
}

__device__ __forceinline__ bool sk_row(const SkFamRows &F, int64_t r, cd *q, int *f) {
  const int e = (int)(r / F.n);
  if (e >= 4 || F.fam[e] < 0) return false;
  const cd v = F.Q[r];
  if (!(isfin(v)) || (v.r == 0.0 && v.i == 0.0)) return false;
  *q = v;
  *f = F.fam[e];
  return true;
}

constexpr int SK_CH = 64;

// Householder QR of the (N + SK_CH) x N matrix A (LDS, column-major with
// leading dimension LD) by 256 threads; afterwards rows 0..N-1 hold R.
template <int NMAX>
__device__ void sk_householder(cd *A, int N, cd *red, double *vvs, cd *v0s) {
  constexpr int LD = NMAX + SK_CH;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int M = N + SK_CH;
  for (int k = 0; k < N; ++k) {
    // norm of column k below the diagonal
    double s = 0.0;
    for (int i = k + tid; i < M; i += 256) s += abs2(A[i + LD * k]);
    red[tid].r = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (tid < o) red[tid].r += red[tid + o].r;
      __syncthreads();
    }
    const double nrm2 = red[0].r;
    const cd alpha = A[k + LD * k];
    __syncthreads();  // every thread has read red[0] before thread 0 reuses it
    if (tid == 0) {
      const double nx = sqrt(nrm2), aa = cabs(alpha);
      // beta = -exp(i arg alpha) |x|, v = x - beta e1, v^H v = 2 |x| (|x| + |alpha|)
      const cd ph = aa > 0.0 ? cd{alpha.r / aa, alpha.i / aa} : cd{1.0, 0.0};
      const cd beta = cd{-ph.r * nx, -ph.i * nx};
      *v0s = alpha - beta;
      *vvs = 2.0 * nx * (nx + aa);
      red[0] = beta;
    }
    __syncthreads();
    const double vv = *vvs;
    const cd v0 = *v0s, beta = red[0];
    __syncthreads();
    if (nrm2 < 1e-290) {
      // every entry of the column below 1e-145: its squares underflow and
      // 2/vv overflows, so the reflection would be garbage.  The entries are
      // dropped (zero), far below any interpolation tolerance.
      for (int i = k + 1 + tid; i < M; i += 256) A[i + LD * k] = cd{0.0, 0.0};
    } else if (vv > 0.0) {
      // columns j > k: w = v^H A[k:, j]; A[k:, j] -= (2/vv) v w, one wave per column
      for (int j = k + 1 + wave; j < N; j += 4) {
        cd w = cd{0.0, 0.0};
        for (int i = k + lane; i < M; i += 64) {
          const cd vi = i == k ? v0 : A[i + LD * k];
          w = fmac_cj(w, vi, A[i + LD * j]);
        }
        for (int o = 32; o > 0; o >>= 1) {
          w.r += __shfl_xor(w.r, o);
          w.i += __shfl_xor(w.i, o);
        }
        const cd g = rmul(2.0 / vv, w);
        for (int i = k + lane; i < M; i += 64) {
          const cd vi = i == k ? v0 : A[i + LD * k];
          A[i + LD * j] = A[i + LD * j] - vi * g;
        }
      }
      __syncthreads();
      for (int i = k + tid; i < M; i += 256) A[i + LD * k] = i == k ? beta : cd{0.0, 0.0};
    }
    __syncthreads();
  }
}

// mode 0: rows from the family F (nrows = 4 n); mode 1: rows of the stacked
// partial factors Rin (nrows = parts * N, Rin[part][N x N] column-major).
template <int NMAX>
__global__ __launch_bounds__(256) void k_sk_tsqr(SkFamRows F, const cd *__restrict__ Rin, int mode, int N,
                                                 int64_t nrows, cd *__restrict__ Rout) {
  constexpr int LD = NMAX + SK_CH;
  __shared__ cd A[LD * NMAX];
  __shared__ cd red[256];
  __shared__ double vvs;
  __shared__ cd v0s;
  const int tid = threadIdx.x;
  for (int e = tid; e < LD * NMAX; e += 256) A[e] = cd{0.0, 0.0};
  __syncthreads();
  const int64_t nch = (nrows + SK_CH - 1) / SK_CH;
  for (int64_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
    for (int e = tid; e < SK_CH * N; e += 256) {
      const int i = e % SK_CH, j = e / SK_CH;
      const int64_t r = ch * SK_CH + i;
      cd a = cd{0.0, 0.0};
      if (r < nrows) {
        if (mode == 0) {
          cd q;
          int f;
          if (sk_row(F, r, &q, &f)) a = sk_exp(q, F.sig[f] * F.z[j] + F.c[f]);
        } else {
          const int64_t part = r / N, row = r % N;
          a = Rin[part * N * N + row + (int64_t)N * j];
        }
      }
      A[N + i + LD * j] = a;
    }
    __syncthreads();
    sk_householder<NMAX>(A, N, red, &vvs, &v0s);
  }
  for (int e = tid; e < N * N; e += 256) {
    const int i = e % N, j = e / N;
    Rout[(int64_t)blockIdx.x * N * N + e] = i <= j ? A[i + LD * j] : cd{0.0, 0.0};
  }
}

// Interpolation error: one thread per (row, plane).  P (N x k) column-major,
// idx (k).  err holds the bit pattern of the largest |residual| (NaN -> Inf).
__global__ void k_sk_err(SkFamRows F, int64_t nrows, int N, int k, const int *__restrict__ idx,
                         const cd *__restrict__ P, unsigned long long *err) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nrows * N) return;
  const int64_t r = t / N;
  const int j = (int)(t % N);
  cd q;
  int f;
  if (!sk_row(F, r, &q, &f)) return;
  const double s = F.sig[f], c = F.c[f];
  cd res = sk_exp(q, s * F.z[j] + c);
  for (int m = 0; m < k; ++m) res = res - sk_exp(q, s * F.z[idx[m]] + c) * P[j + (int64_t)N * m];
  double e = cabs(res);
  if (!(e <= 1.7976931348623157e308)) e = __builtin_huge_val();
  atomicMax(err, (unsigned long long)__double_as_longlong(e));
}

__global__ void k_sk_zero(unsigned long long *err) {
  if (threadIdx.x == 0) *err = 0ull;
}

template <int NMAX>
static hipError_t tsqr(const SkFamRows &F, const cd *Rin, int mode, int N, int64_t nrows, cd *Rout, int blocks,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_sk_tsqr<NMAX>, dim3(blocks), dim3(256), 0, s, F, Rin, mode, N, nrows, Rout);
  return hipGetLastError();
}

hipError_t launch_sk_tsqr(const SkFamRows &F, const cd *Rin, int mode, int N, int64_t nrows, cd *Rout, int blocks,
                          hipStream_t s) {
  if (N <= 16) return tsqr<16>(F, Rin, mode, N, nrows, Rout, blocks, s);
  if (N <= 32) return tsqr<32>(F, Rin, mode, N, nrows, Rout, blocks, s);
  if (N <= 64) return tsqr<64>(F, Rin, mode, N, nrows, Rout, blocks, s);
  return hipErrorInvalidValue;
}

hipError_t launch_sk_err(const SkFamRows &F, int64_t nrows, int N, int k, const int *idx, const cd *P,
                         unsigned long long *err, hipStream_t s) {
  hipLaunchKernelGGL(k_sk_zero, dim3(1), dim3(64), 0, s, err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t nt = nrows * N;
  if (nt == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sk_err, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, F, nrows, N, k, idx, P, err);
  return hipGetLastError();
}

}  // namespace ytf
