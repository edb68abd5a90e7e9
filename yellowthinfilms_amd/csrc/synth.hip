// synth.hip -- radial Bessel synthesis of the 1D (azimuthally symmetric)
// closed-form Green's function, FP64.
//
// Reference: closed_form_G for a CavityTransferCache1D
// (mod_ClosedFormField.jl:435-487), the pixel integral _integrate_pixel_1d
// (:344-390) over the Bessel row cache _build_bessel_row_cache_vol (:304-319),
// with the weights of _weighted_mcache (:287-294).
//
// For a pixel on the local x axis at radius rho the reference sums, over
// nodes k (kx_k, ky = 0) and the two modes of each direction,
//   I = (i/2)(J2-J0) A E_x - (i/2)(J0+J2) B E_y + J1 C E_x   (x component, etc.)
// with A, B, C the moment projections of the weighted amplitudes.  Summing the
// modes first, every term is an entry of the node's spectral dyadic
// g~(kx_k, 0; z, z') (the per-node strength divided by its weight, the output
// of the node kernel).  Per depth pair the 3x3 response is then
//   R(rho) = sum_k w'_k [ J0(kx_k rho) a0_k + J1(kx_k rho) a1_k + J2(kx_k rho) a2_k ],
//   w'_k = kx_k * i wt_k / (2 pi),
// with nine independent complex coefficient columns per node:
//   J0: a = -(i/2)(gxx + gyy),  b = -(i/2)(gxy - gyx),  e = -i gzz
//   J1: gxz, gyz, gzx, gzy
//   J2: c =  (i/2)(gxx - gyy),  d =  (i/2)(gxy + gyx)
// and R_xx = a + c, R_yx = -b + d, R_xy = b + d, R_yy = a - c, R_zz = e,
// R_xz = gxz-column, R_yz, R_zx, R_zy likewise (g entry c + 3q).
#include <hip/hip_runtime.h>

#include "ytf_kernels.h"

namespace ytf {

// column order of the coefficient arrays
enum { SC_A = 0, SC_B, SC_E, SC_XZ, SC_YZ, SC_ZX, SC_ZY, SC_C, SC_D, SC_N };

// J0, J1, J2 at x (J1 odd, J0 and J2 even).  J0 and J1 are the device
// library's; J2 = 2 J1(x)/x - J0(x), or its power series below |x| = 1/2
// where that difference cancels.
__device__ inline void bessel012(double x, double &b0, double &b1, double &b2) {
  const double ax = fabs(x);
  b0 = __ocml_j0_f64(ax);
  const double t = __ocml_j1_f64(ax);
  b1 = x < 0.0 ? -t : t;
  if (ax < 0.5) {
    // J2(x) = sum_m (-1)^m (x/2)^(2m+2) / (m! (m+2)!)
    const double h2 = 0.25 * ax * ax;
    double term = 0.5 * h2, s = term;
#pragma unroll
    for (int m = 1; m < 9; ++m) {
      term *= -h2 / (double)(m * (m + 2));
      s += term;
    }
    b2 = s;
  } else {
    b2 = 2.0 * t / ax - b0;
  }
}

// g~[p][9][n] -> coefficient columns coef[p][SC_N][n] (complex), weight and kx folded in.
__global__ void k_synth_coef(int64_t n, int np, const double *__restrict__ kx, const double *__restrict__ wt,
                             const cd *__restrict__ g, cd *__restrict__ coef) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (k >= n || p >= np) return;
  const cd *gp = g + (size_t)p * 9 * n;
  const cd gxx = gp[0 * n + k], gyx = gp[1 * n + k], gzx = gp[2 * n + k];
  const cd gxy = gp[3 * n + k], gyy = gp[4 * n + k], gzy = gp[5 * n + k];
  const cd gxz = gp[6 * n + k], gyz = gp[7 * n + k], gzz = gp[8 * n + k];
  // w' = kx * i wt / (2 pi) = i wi, so -(i/2) w' = wi/2, (i/2) w' = -wi/2, -i w' = wi
  const double wi = kx[k] * wt[k] * 0.15915494309189533577;
  const double hw = 0.5 * wi;
  cd *o = coef + (size_t)p * SC_N * n;
  o[SC_A * n + k] = rmul(hw, gxx + gyy);
  o[SC_B * n + k] = rmul(hw, gxy - gyx);
  o[SC_E * n + k] = rmul(wi, gzz);
  o[SC_XZ * n + k] = times_i(rmul(wi, gxz));
  o[SC_YZ * n + k] = times_i(rmul(wi, gyz));
  o[SC_ZX * n + k] = times_i(rmul(wi, gzx));
  o[SC_ZY * n + k] = times_i(rmul(wi, gzy));
  o[SC_C * n + k] = rmul(-hw, gxx - gyy);
  o[SC_D * n + k] = rmul(-hw, gxy + gyx);
}

// Partial contraction over one node chunk.  Thread = one radius; the block's
// PC depth pairs share the Bessel values.  part[ks][p][SC_N][nr].
template <int PC>
__global__ __launch_bounds__(256) void k_synth_fp64(int64_t n, int np, int nr, int64_t chunk,
                                                    const double *__restrict__ kx, const double *__restrict__ rho,
                                                    const cd *__restrict__ coef, cd *__restrict__ part) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int p0 = blockIdx.y * PC;
  const int ks = blockIdx.z;
  const int64_t k0 = ks * chunk;
  int64_t k1 = k0 + chunk;
  if (k1 > n) k1 = n;
  const double rh = r < nr ? rho[r] : 0.0;
  cd acc[PC][SC_N];
#pragma unroll
  for (int q = 0; q < PC; ++q)
#pragma unroll
    for (int i = 0; i < SC_N; ++i) acc[q][i] = cd{0.0, 0.0};
  for (int64_t k = k0; k < k1; ++k) {
    double b0, b1, b2;
    bessel012(kx[k] * rh, b0, b1, b2);
#pragma unroll
    for (int q = 0; q < PC; ++q) {
      if (p0 + q >= np) break;
      const cd *c = coef + (size_t)(p0 + q) * SC_N * n + k;
#pragma unroll
      for (int i = 0; i < SC_N; ++i) {
        const double b = i < SC_XZ ? b0 : (i < SC_C ? b1 : b2);
        const cd v = c[(size_t)i * n];
        acc[q][i].r = fma(b, v.r, acc[q][i].r);
        acc[q][i].i = fma(b, v.i, acc[q][i].i);
      }
    }
  }
  if (r >= nr) return;
#pragma unroll
  for (int q = 0; q < PC; ++q) {
    if (p0 + q >= np) break;
    cd *o = part + ((size_t)ks * np + p0 + q) * SC_N * nr + r;
#pragma unroll
    for (int i = 0; i < SC_N; ++i) o[(size_t)i * nr] = acc[q][i];
  }
}

// Sum the node-chunk partials in a fixed order and form the nine entries:
// R[p][c + 3q][r].
__global__ void k_synth_reduce(int np, int nr, int nks, const cd *__restrict__ part, cd *__restrict__ R) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (r >= nr) return;
  cd s[SC_N];
#pragma unroll
  for (int i = 0; i < SC_N; ++i) s[i] = cd{0.0, 0.0};
  for (int ks = 0; ks < nks; ++ks) {
    const cd *q = part + ((size_t)ks * np + p) * SC_N * nr + r;
#pragma unroll
    for (int i = 0; i < SC_N; ++i) s[i] = s[i] + q[(size_t)i * nr];
  }
  cd *o = R + (size_t)p * 9 * nr + r;
  o[0 * nr] = s[SC_A] + s[SC_C];   // xx
  o[1 * nr] = s[SC_D] - s[SC_B];   // yx
  o[2 * nr] = s[SC_ZX];            // zx
  o[3 * nr] = s[SC_B] + s[SC_D];   // xy
  o[4 * nr] = s[SC_A] - s[SC_C];   // yy
  o[5 * nr] = s[SC_ZY];            // zy
  o[6 * nr] = s[SC_XZ];            // xz
  o[7 * nr] = s[SC_YZ];            // yz
  o[8 * nr] = s[SC_E];             // zz
}

// G(i, j, p, c, q) = (Rot R(rho_ij) Rot^T)[c, q], Rot the in-plane rotation to
// the pixel's azimuth (mod_ClosedFormField.jl:474-484).
__global__ void k_synth_rotate(int nx, int ny, int np, int nr, const double *__restrict__ xs,
                               const double *__restrict__ ys, const int *__restrict__ rmap,
                               const cd *__restrict__ R, cd *__restrict__ G) {
  const int64_t pix = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  const int64_t npix = (int64_t)nx * ny;
  if (pix >= npix) return;
  const int i = (int)(pix % nx), j = (int)(pix / nx);
  const double x = xs[i], y = ys[j];
  const double rr = hypot(x, y);
  const double cp = rr == 0.0 ? 1.0 : x / rr, sp = rr == 0.0 ? 0.0 : y / rr;
  const double Rot[3][3] = {{cp, -sp, 0.0}, {sp, cp, 0.0}, {0.0, 0.0, 1.0}};  // Rot[row][col]
  const cd *Rr = R + (size_t)p * 9 * nr + rmap[pix];
  cd M[3][3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) M[c][q] = Rr[(size_t)(c + 3 * q) * nr];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      cd s{0.0, 0.0};
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) s = s + rmul(Rot[c][a] * Rot[q][b], M[a][b]);
      G[pix + npix * ((size_t)p + (size_t)np * (c + 3 * q))] = s;
    }
}

hipError_t launch_synth_coef(int64_t n, int np, const double *kx, const double *wt, const cd *g, cd *coef,
                             hipStream_t s) {
  dim3 grid((unsigned)((n + 255) / 256), np);
  hipLaunchKernelGGL(k_synth_coef, grid, dim3(256), 0, s, n, np, kx, wt, g, coef);
  return hipGetLastError();
}

int synth_fp64_splits(int64_t n, int np, int nr) {
  const int PC = 4;
  const int64_t waves = (int64_t)((nr + 63) / 64) * ((np + PC - 1) / PC);
  int64_t ks = (4096 + waves - 1) / waves;
  int64_t maxks = (n + 31) / 32;
  if (ks > maxks) ks = maxks;
  if (ks < 1) ks = 1;
  return (int)ks;
}

hipError_t launch_synth_fp64(int64_t n, int np, int nr, int nks, const double *kx, const double *rho,
                             const cd *coef, cd *part, hipStream_t s) {
  const int PC = 4;
  const int64_t chunk = (n + nks - 1) / nks;
  dim3 grid((unsigned)((nr + 255) / 256), (unsigned)((np + PC - 1) / PC), (unsigned)nks);
  hipLaunchKernelGGL(k_synth_fp64<PC>, grid, dim3(256), 0, s, n, np, nr, chunk, kx, rho, coef, part);
  return hipGetLastError();
}

hipError_t launch_synth_reduce(int np, int nr, int nks, const cd *part, cd *R, hipStream_t s) {
  dim3 g2((unsigned)((nr + 255) / 256), np);
  hipLaunchKernelGGL(k_synth_reduce, g2, dim3(256), 0, s, np, nr, nks, part, R);
  return hipGetLastError();
}

hipError_t launch_synth_rotate(int nx, int ny, int np, int nr, const double *xs, const double *ys, const int *rmap,
                               const cd *R, cd *G, hipStream_t s) {
  const int64_t npix = (int64_t)nx * ny;
  dim3 grid((unsigned)((npix + 255) / 256), np);
  hipLaunchKernelGGL(k_synth_rotate, grid, dim3(256), 0, s, nx, ny, np, nr, xs, ys, rmap, R, G);
  return hipGetLastError();
}

}  // namespace ytf
