// ktable.hip -- the tabulated k_rho surrogate of the scattered spectral dyadic
// of a stack that is invariant under rotations about z, evaluated on any
// (kx, ky) node set (no counterpart in the reference, which solves every node).
//
// With source and observation point in layer 1 above the stack,
//     g~_s(kx, ky) = R(phi) g~_s(k_rho, 0) R(phi)^T,
//     g~_s(k_rho, 0) = P omega (1 + k_rho^2 / omega^2) / kz0 * t(k_rho),
// P = exp(i kz0 zsum), kz0 = sqrt(omega^2 - k_rho^2) (Im >= 0), R the rotation
// about z by phi = atan2(ky, kx).  The table holds t (9 complex entries,
// c + 3q) on contiguous k_rho panels, each sampled on a uniform grid of m_p
// points in a panel variable s that keeps t analytic across a branch point
// k_b on the real axis (the air light line, or that of a lossless substrate):
//     kind 0: s = k_rho;  kind 1: s = sqrt(k_b^2 - k_rho^2) (k_rho <= k_b);
//     kind 2: s = sqrt(k_rho^2 - k_b^2) (k_rho >= k_b).
// Kind bit 4 marks a panel that did not reach the table's tolerance: its nodes
// are flagged for the exact solve, as are nodes outside the table and nodes
// on the air light line itself (kz0 = 0).
//
// One thread per node: binary search of the panel boundaries (staged in LDS),
// four-point (cubic) Lagrange interpolation of the 18 reals on the panel's
// grid, the exact factors, the rotation, and nine coalesced complex stores
// ([9][n], entry-major); flagged rows are appended to a row list with one
// atomic per wave.  The table is entry-major (entry e of row r at
// val[e * nrow + r]): a wave's lanes sit on nearby rows, so each of the 72
// loads of a node touches a few cache lines for the whole wave, where the
// row-major table (a 144-byte row per lane) touched one line per lane and the
// L1's line rate bounded the kernel (61 -> 45 us per 2^20 nodes).  A
// Chebyshev series per panel, evaluated by Clenshaw, was measured first:
// every lane then reads its panel's 17 x 18 coefficients (~0.08 ms).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ytf_kernels.h"

namespace ytf {

__global__ __launch_bounds__(256) void k_ktable_ghat(int64_t n, const double *__restrict__ kx,
                                                     const double *__restrict__ ky, double w, double zsum, KTable T,
                                                     cd *__restrict__ ghat, unsigned char *__restrict__ flags,
                                                     int *__restrict__ count, int *__restrict__ list) {
  extern __shared__ double kbs[];  // T.np + 1 panel boundaries, then T.np grid scales
  double *invh = kbs + T.np + 1;
  for (int i = threadIdx.x; i <= T.np; i += blockDim.x) kbs[i] = T.kb[i];
  for (int q = threadIdx.x; q < T.np; q += blockDim.x)
    invh[q] = (double)(T.m[q] - 1) / (T.sab[3 * q + 1] - T.sab[3 * q]);
  __syncthreads();
  const int64_t nrow = T.row0[T.np - 1] + T.m[T.np - 1];  // entry e of row r at val[e * nrow + r]
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double a = 0.0, b = 0.0, kr = 0.0;
  int p = -1;
  if (i < n) {
    a = kx[i];
    b = ky[i];
    kr = sqrt(a * a + b * b);
    if (kr >= kbs[0] && kr <= kbs[T.np]) {
      int lo = 0, hi = T.np;  // kbs[lo] <= kr < kbs[hi] (kr == kbs[np] falls in the last panel)
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (kbs[mid] <= kr) lo = mid;
        else hi = mid;
      }
      p = lo;
    }
  }
  // per lane: panel kind, kz0, flag, and the interpolated t
  const int kind = (i < n && p >= 0) ? T.kind[p] : 4;
  const double d = (w - kr) * (w + kr);  // exactly 0 on the light line, accurate near it
  const cd kz = d >= 0.0 ? cd{sqrt(d), 0.0} : cd{0.0, sqrt(-d)};
  const bool flag = i < n && ((kind & 4) != 0 || (kz.r == 0.0 && kz.i == 0.0));
  const bool work = i < n && !flag;
  double val[18];
  if (work) {
    const double s_lo = T.sab[3 * p], kb = T.sab[3 * p + 2], inv_h = invh[p];
    const double sv = kind == 0 ? kr : sqrt(fmax(kind == 1 ? kb * kb - kr * kr : kr * kr - kb * kb, 0.0));
    const int m = T.m[p];
    const double u = fmin((double)(m - 1), fmax(0.0, (sv - s_lo) * inv_h));
    const int j = min(max((int)u, 1), m - 3);  // rows j - 1 .. j + 2
    const double r = u - j;
    const double rm = r - 1.0, rp = r + 1.0, r2 = r - 2.0;
    const double w0 = -r * rm * r2 * (1.0 / 6.0), w1 = 0.5 * rp * rm * r2, w2 = -0.5 * rp * r * r2,
                 w3 = rp * r * rm * (1.0 / 6.0);
    const double *v = T.val + (T.row0[p] + j - 1);
#pragma unroll
    for (int e = 0; e < 18; ++e) {
      const double *ve = v + e * nrow;
      val[e] = __builtin_fma(w0, ve[0], __builtin_fma(w1, ve[1], __builtin_fma(w2, ve[2], w3 * ve[3])));
    }
  }
  if (i < n) {
    if (work) {
      cd G[9];
      // F = omega (1 + kr^2 / omega^2) / kz0 * exp(i kz0 zsum), kz0 real or i kappa
      const double grow = __builtin_fma(kr * kr, T.inv_w, w);
      const double kzm = kz.r + kz.i;  // |kz0|: one of the two is 0
      const double m = grow / kzm;
      cd F;
      if (kz.i == 0.0) {  // propagating
        double sn, cs;
        sincos(kzm * zsum, &sn, &cs);
        F = cd{m * cs, m * sn};
      } else {  // evanescent: 1/kz0 = -i / kappa
        F = cd{0.0, -m * exp(-kzm * zsum)};
      }
#pragma unroll
      for (int e = 0; e < 9; ++e) G[e] = F * cd{val[e], val[9 + e]};
      // rotation: H = G R^T, then R H (entry a + 3b is row a, column b)
      const double ikr = kr > 0.0 ? 1.0 / kr : 0.0;
      const double c = kr > 0.0 ? a * ikr : 1.0, s = b * ikr;
      cd H[9];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const cd gx = G[r], gy = G[r + 3];
        H[r] = rmul(c, gx) - rmul(s, gy);
        H[r + 3] = rmul(s, gx) + rmul(c, gy);
        H[r + 6] = G[r + 6];
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const cd hx = H[3 * q], hy = H[3 * q + 1];
        ghat[(int64_t)(3 * q) * n + i] = rmul(c, hx) - rmul(s, hy);
        ghat[(int64_t)(3 * q + 1) * n + i] = rmul(s, hx) + rmul(c, hy);
        ghat[(int64_t)(3 * q + 2) * n + i] = H[3 * q + 2];
      }
    }
    flags[i] = flag ? 1 : 0;
  }
  const unsigned long long bal = __ballot(flag);
  if (bal == 0) return;
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0) base = atomicAdd(count, __popcll(bal));
  base = __shfl(base, 0);
  if (flag) list[base + __popcll(bal & ((1ull << lane) - 1ull))] = (int)i;
}

hipError_t launch_ktable_ghat(int64_t n, const double *kx, const double *ky, double w, double zsum, const KTable &T,
                              cd *ghat, unsigned char *flags, int *count, int *list, hipStream_t s) {
  if (n == 0) return hipSuccess;
  KTable U = T;
  U.inv_w = 1.0 / w;
  hipLaunchKernelGGL(k_ktable_ghat, dim3((unsigned)((n + 255) / 256)), dim3(256), sizeof(double) * (2 * T.np + 1), s,
                     n, kx, ky, w, zsum, U, ghat, flags, count, list);
  return hipGetLastError();
}

}  // namespace ytf
