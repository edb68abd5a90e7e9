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

#include <algorithm>

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

// Per-node terms of the radial synthesis: T[k][p][c + 3q][r] = node k's
// term of R_p(rho_r), i.e. R_p(rho_r) of a one-node cache (kx_k, wt_k).  One
// thread per (node, radius), depth pair p = blockIdx.y.
__global__ void k_synth_terms(int64_t n, int np, int nr, const double *__restrict__ kx, const double *__restrict__ wt,
                              const cd *__restrict__ g, const double *__restrict__ rho, cd *__restrict__ T) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (e >= n * nr || p >= np) return;
  const int64_t k = e / nr;
  const int r = (int)(e - k * nr);
  const cd *gp = g + (size_t)p * 9 * n;
  const cd gxx = gp[0 * n + k], gyx = gp[1 * n + k], gzx = gp[2 * n + k];
  const cd gxy = gp[3 * n + k], gyy = gp[4 * n + k], gzy = gp[5 * n + k];
  const cd gxz = gp[6 * n + k], gyz = gp[7 * n + k], gzz = gp[8 * n + k];
  const double wi = kx[k] * wt[k] * 0.15915494309189533577;
  const double hw = 0.5 * wi;
  double b0, b1, b2;
  bessel012(kx[k] * rho[r], b0, b1, b2);
  const cd a = rmul(b0 * hw, gxx + gyy), b = rmul(b0 * hw, gxy - gyx);
  const cd c = rmul(-b2 * hw, gxx - gyy), d = rmul(-b2 * hw, gxy + gyx);
  cd *o = T + ((size_t)k * np + p) * 9 * nr + r;
  o[0 * nr] = a + c;                              // xx
  o[1 * nr] = d - b;                              // yx
  o[2 * nr] = times_i(rmul(b1 * wi, gzx));        // zx
  o[3 * nr] = b + d;                              // xy
  o[4 * nr] = a - c;                              // yy
  o[5 * nr] = times_i(rmul(b1 * wi, gzy));        // zy
  o[6 * nr] = times_i(rmul(b1 * wi, gxz));        // xz
  o[7 * nr] = times_i(rmul(b1 * wi, gyz));        // yz
  o[8 * nr] = rmul(b0 * wi, gzz);                 // zz
}

// FP64 path: a Bessel table of one batch of nodes, then a contraction per
// node chunk into partials, a fixed-order sum over the chunks, and the nine
// entries.
//
// J[ord][kk][r] = J_ord(kx[k0 + kk] rho[r]), ord = 0, 1, 2 (r fastest).
__global__ void k_synth_bessel(int64_t k0, int64_t nk, int nr, const double *__restrict__ kx,
                               const double *__restrict__ rho, double *__restrict__ J) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nk * nr) return;
  const int64_t kk = e / nr;
  const int r = (int)(e - kk * nr);
  double b0, b1, b2;
  bessel012(kx[k0 + kk] * rho[r], b0, b1, b2);
  const int64_t ne = nk * nr;
  J[e] = b0;
  J[ne + e] = b1;
  J[2 * ne + e] = b2;
}

// Contraction of node chunk c (nodes [c chunk, min((c + 1) chunk, n)), inside
// the table's batch starting at node k0): part[c][p][SC_N][nr].  Lane = radius
// (64 per block); wave w of the block takes the depth pairs 8 blockIdx.y + 2w
// and + 1.
__global__ __launch_bounds__(256) void k_synth_contract(int64_t n, int np, int nr, int64_t chunk, int c0, int64_t k0,
                                                        int64_t nk, const double *__restrict__ J,
                                                        const cd *__restrict__ coef, cd *__restrict__ part) {
  const int r = blockIdx.x * 64 + (threadIdx.x & 63);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p0 = blockIdx.y * 8 + 2 * wv;
  const int c = c0 + blockIdx.z;
  const int64_t ka = c * chunk, kb = min(ka + chunk, n);
  const bool live = r < nr;
  const int64_t ne = nk * nr;
  cd acc[2][SC_N];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int i = 0; i < SC_N; ++i) acc[q][i] = cd{0.0, 0.0};
  const int nq = min(2, np - p0);
  // Per step of CP nodes the block stages in LDS the table values (wave u
  // loads node kg + u for the block's 64 radii) and the coefficients of its
  // 8 depth pairs; the next step's values are in flight while the current CP
  // nodes are summed.  Coefficient reads are wave-uniform (LDS broadcast).
  constexpr int CP = 4, NROW = 8 * SC_N;  // coefficient rows (pair, column) per block
  __shared__ double jl[2][CP][3][64];
  __shared__ cd cl[2][NROW][CP];
  const int lane = threadIdx.x & 63, pb = blockIdx.y * 8;
  auto gload = [&](int64_t kg, double *b, cd *cv) {
    const int64_t k = kg + wv;
    const bool ok = live && k < kb;
    const int64_t e = (k - k0) * nr + r;
#pragma unroll
    for (int o = 0; o < 3; ++o) b[o] = ok ? J[o * ne + e] : 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int x = threadIdx.x + 256 * h;  // coefficient element (row, node)
      const int row = x / CP, u = x % CP;
      const int p = pb + row / SC_N;
      const bool okc = x < NROW * CP && p < np && kg + u < kb;
      cv[h] = okc ? coef[((size_t)p * SC_N + row % SC_N) * n + kg + u] : cd{0.0, 0.0};
    }
  };
  auto stash = [&](int b2, const double *b, const cd *cv) {
#pragma unroll
    for (int o = 0; o < 3; ++o) jl[b2][wv][o][lane] = b[o];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int x = threadIdx.x + 256 * h;
      if (x < NROW * CP) cl[b2][x / CP][x % CP] = cv[h];
    }
  };
  double g[3];
  cd gc[2];
  gload(ka, g, gc);
  stash(0, g, gc);
  __syncthreads();
  int buf = 0;
  for (int64_t kg = ka; kg < kb; kg += CP, buf ^= 1) {
    gload(kg + CP, g, gc);
#pragma unroll
    for (int u = 0; u < CP; ++u) {
      if (kg + u < kb) {
        const double b0 = jl[buf][u][0][lane], b1 = jl[buf][u][1][lane], b2 = jl[buf][u][2][lane];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (q < nq) {
#pragma unroll
            for (int i = 0; i < SC_N; ++i) {
              const double b = i < SC_XZ ? b0 : (i < SC_C ? b1 : b2);
              const cd v = cl[buf][(2 * wv + q) * SC_N + i][u];
              acc[q][i].r = fma(b, v.r, acc[q][i].r);
              acc[q][i].i = fma(b, v.i, acc[q][i].i);
            }
          }
        }
      }
    }
    stash(buf ^ 1, g, gc);
    __syncthreads();
  }
  if (!live) return;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    if (q < nq) {
      cd *o = part + ((size_t)c * np + p0 + q) * SC_N * nr + r;
#pragma unroll
      for (int i = 0; i < SC_N; ++i) o[(size_t)i * nr] = acc[q][i];
    }
  }
}

// S[p][i][r] = sum over the nc chunks, in chunk order, of part[c][p][i][r].
__global__ void k_synth_chunksum(int64_t m, int nc, const cd *__restrict__ part, cd *__restrict__ S) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= m) return;
  cd s{0.0, 0.0};
  int c = 0;
  for (; c + 8 <= nc; c += 8) {
    cd v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part[(size_t)(c + u) * m + e];
#pragma unroll
    for (int u = 0; u < 8; ++u) s = s + v[u];
  }
  for (; c < nc; ++c) s = s + part[(size_t)c * m + e];
  S[e] = s;
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

SynthPlan synth_fp64_plan(int64_t n, int np, int nr) {
  SynthPlan P;
  const int64_t groups = (int64_t)((nr + 63) / 64) * ((np + 7) / 8) * 4;  // waves per chunk
  int64_t C = (2048 + groups - 1) / groups;
  const int64_t cmax = (n + 15) / 16;  // at least 16 nodes per chunk
  if (C > cmax) C = cmax;
  if (C < 1) C = 1;
  int64_t chunk = (n + C - 1) / C;
  const int64_t tab = (int64_t)1 << 23;  // table entries per order and batch
  if (chunk * nr > tab) chunk = std::max<int64_t>(1, tab / std::max(nr, 1));
  if (chunk < 1) chunk = 1;
  P.chunk = chunk;
  P.nchunks = (int)((n + chunk - 1) / chunk);
  P.chunks_per_batch = (int)std::max<int64_t>(1, std::min<int64_t>(P.nchunks, tab / std::max<int64_t>(chunk * nr, 1)));
  P.batch_nodes = std::min<int64_t>(n, chunk * P.chunks_per_batch);
  return P;
}

hipError_t launch_synth_fp64(int64_t n, int np, int nr, const SynthPlan &P, const double *kx, const double *rho,
                             const cd *coef, double *J, cd *part, cd *S, hipStream_t s) {
  for (int c0 = 0; c0 < P.nchunks; c0 += P.chunks_per_batch) {
    const int nc = std::min(P.chunks_per_batch, P.nchunks - c0);
    const int64_t k0 = c0 * P.chunk, nk = std::min<int64_t>((int64_t)nc * P.chunk, n - k0);
    const int64_t ne = nk * nr;
    hipLaunchKernelGGL(k_synth_bessel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, k0, nk, nr, kx, rho, J);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 grid((unsigned)((nr + 63) / 64), (unsigned)((np + 7) / 8), (unsigned)nc);
    hipLaunchKernelGGL(k_synth_contract, grid, dim3(256), 0, s, n, np, nr, P.chunk, c0, k0, nk, J, coef, part);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  const int64_t m = (int64_t)np * SC_N * nr;
  hipLaunchKernelGGL(k_synth_chunksum, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, P.nchunks,
                     (const cd *)part, S);
  return hipGetLastError();
}

hipError_t launch_synth_terms(int64_t n, int np, int nr, const double *kx, const double *wt, const cd *g,
                              const double *rho, cd *T, hipStream_t s) {
  const int64_t m = n * nr;
  if (m == 0 || np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_synth_terms, dim3((unsigned)((m + 255) / 256), (unsigned)np), dim3(256), 0, s, n, np, nr, kx, wt,
                     g, rho, T);
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
