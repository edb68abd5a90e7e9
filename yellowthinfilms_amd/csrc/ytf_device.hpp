// ytf_device.hpp -- FP64 complex arithmetic and fixed-size matrix helpers for
// the gfx950 node kernels.  Everything is written for full unrolling: all
// matrix indices are compile-time constants after inlining, so the matrices
// live in VGPRs; data-dependent choices (pivots, mode slots, swaps) are made
// with selects, never with dynamic register-array indexing.
#pragma once
#include <hip/hip_runtime.h>

namespace ytf {

struct cd {
  double r, i;
};

__device__ __forceinline__ cd mk(double r, double i = 0.0) { return cd{r, i}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cd operator-(cd a) { return cd{-a.r, -a.i}; }
__device__ __forceinline__ cd operator*(cd a, cd b) {
  return cd{__builtin_fma(a.r, b.r, -a.i * b.i), __builtin_fma(a.r, b.i, a.i * b.r)};
}
__device__ __forceinline__ cd rmul(double s, cd a) { return cd{s * a.r, s * a.i}; }
__device__ __forceinline__ cd conj(cd a) { return cd{a.r, -a.i}; }
__device__ __forceinline__ cd times_i(cd a) { return cd{-a.i, a.r}; }
__device__ __forceinline__ double abs2(cd a) { return __builtin_fma(a.r, a.r, a.i * a.i); }
__device__ __forceinline__ double cabs(cd a) { return hypot(a.r, a.i); }
__device__ __forceinline__ double cabs1(cd a) { return fabs(a.r) + fabs(a.i); }
__device__ __forceinline__ bool isfin(cd a) { return isfinite(a.r) && isfinite(a.i); }
// a += b*c
__device__ __forceinline__ cd fmac(cd acc, cd b, cd c) {
  return cd{__builtin_fma(b.r, c.r, __builtin_fma(-b.i, c.i, acc.r)), __builtin_fma(b.r, c.i, __builtin_fma(b.i, c.r, acc.i))};
}
// conj(b)*c + acc
__device__ __forceinline__ cd fmac_cj(cd acc, cd b, cd c) {
  return cd{__builtin_fma(b.r, c.r, __builtin_fma(b.i, c.i, acc.r)), __builtin_fma(b.r, c.i, __builtin_fma(-b.i, c.r, acc.i))};
}
// 1/b with one real reciprocal: b is first scaled by an exact power of two
// so that |b'|^2 lies in [1, 8) (b = 0, Inf or NaN give NaN components, as
// the IEEE-division form of this scaling does).  On that range the hardware
// reciprocal and two Newton steps need no special cases, where a division
// would run the general scaled sequence; the result differs from the
// correctly rounded reciprocal by at most an ulp or so.
__device__ __forceinline__ cd cinv(cd b) {
  const double m = fmax(fabs(b.r), fabs(b.i));
  const int e = (m > 0.0 && isfinite(m)) ? ilogb(m) : 0;
  const double br = ldexp(b.r, -e), bi = ldexp(b.i, -e);
  const double den = __builtin_fma(br, br, bi * bi);
  double rden = __builtin_amdgcn_rcp(den);
  rden = __builtin_fma(rden, __builtin_fma(-den, rden, 1.0), rden);
  rden = __builtin_fma(rden, __builtin_fma(-den, rden, 1.0), rden);
  const double sc = ldexp(rden, -e);
  return cd{br * sc, -bi * sc};
}
__device__ __forceinline__ cd operator/(cd a, cd b) { return a * cinv(b); }
// 1/x for a finite nonzero x whose reciprocal is a normal double: the
// hardware reciprocal and two Newton steps (within an ulp or so of 1/x).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}
__device__ __forceinline__ cd cexp(cd z) {
  double e = exp(z.r);
  double s, c;
  sincos(z.i, &s, &c);
  return cd{e * c, e * s};
}
// exp(i*q*d) for real d
__device__ __forceinline__ cd expi(cd q, double d) { return cexp(cd{-q.i * d, q.r * d}); }
// principal square root with C99 signed-zero semantics
__device__ __forceinline__ cd csqrt(cd z) {
  if (z.r == 0.0 && z.i == 0.0) return cd{0.0, z.i};
  double t = sqrt((fabs(z.r) + hypot(z.r, z.i)) * 0.5);
  const double h = t < 1e300 ? 0.5 * rcp_nr(t) : 0.5 / t;  // t > 0 here
  if (z.r >= 0.0) return cd{t, z.i * h};
  return cd{fabs(z.i) * h, copysign(t, z.i)};
}
__device__ __forceinline__ cd sel(bool c, cd a, cd b) { return cd{c ? a.r : b.r, c ? a.i : b.i}; }
// |x| < 1e-250 -> 0 (denormal floor of the stack walk)
__device__ __forceinline__ cd flushd(cd x) { return cabs(x) < 1e-250 ? cd{0.0, 0.0} : x; }

// ---------------------------------------------------------------------------
// Column-major fixed-size matrices: A[i + M*j]
// ---------------------------------------------------------------------------
template <int M, int K, int N>
__device__ __forceinline__ void mm(const cd *A, const cd *B, cd *C) {
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = 0; i < M; ++i) {
      cd s = A[i] * B[K * j];
#pragma unroll
      for (int k = 1; k < K; ++k) s = fmac(s, A[i + M * k], B[k + K * j]);
      C[i + M * j] = s;
    }
}
// 2x2 sub-block (bi,bj) of a 4x4
__device__ __forceinline__ void blk(const cd *S, int bi, int bj, cd *B) {
  B[0] = S[(2 * bi + 0) + 4 * (2 * bj + 0)];
  B[1] = S[(2 * bi + 1) + 4 * (2 * bj + 0)];
  B[2] = S[(2 * bi + 0) + 4 * (2 * bj + 1)];
  B[3] = S[(2 * bi + 1) + 4 * (2 * bj + 1)];
}
__device__ __forceinline__ cd det2(const cd *A) { return A[0] * A[3] - A[2] * A[1]; }
// X = A \ B for 2x2 A, B 2xN (Cramer per column)
template <int N>
__device__ __forceinline__ void solve2(const cd *A, const cd *B, cd *X) {
  const cd rd = cinv(det2(A));
#pragma unroll
  for (int j = 0; j < N; ++j) {
    cd b0 = B[2 * j], b1 = B[2 * j + 1];
    X[2 * j] = (A[3] * b0 - A[2] * b1) * rd;
    X[2 * j + 1] = (A[0] * b1 - A[1] * b0) * rd;
  }
}

}  // namespace ytf
