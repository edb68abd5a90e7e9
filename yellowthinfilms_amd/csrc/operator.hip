// operator.hip -- the 2D lateral sum of closed_form_G and the padded,
// z-structured Green operator on gfx950, FP64.
//
// Lateral sum (closed_form_G for a CavityTransferCache2D,
// mod_ClosedFormField.jl:711-782): per depth pair p and entry e = c + 3q,
//     G(x_i, y_j) = sum_n W_n g~_n exp(i (kx_n x_i + ky_n y_j)),
// either exactly (k_lat_direct, the separable sum of :743-753) or through the
// type-1 NUFFT (k_lat_points + k_lat_strengths feed nufft.hip, :755-780).
//
// Operator (mod_GreenOperator.jl, mod_ClosedFormAssembly.jl): G is sampled on
// the 2Nx x 2Ny padded offsets, FFT'd once (sign (-1)^(x+y) for the centring
// shift, 1/(Px Py) folded in) and stored per plane as [entry][node] with the
// spectral node index y + Py*x (y fastest).  The apply is
//   1. k_opx_fwd: zero-padded FFT along x of each source-plane component,
//      written transposed ([plane][x][y]);
//   2. k_opy<.., false>: FFT along y in place (rows >= Ny are zeros);
//   3. k_op_mix / k_op_zcomb: per spectral node, the z coupling of each block
//      (Toeplitz, Hankel, dense plane pairs, skeleton fold and expansion);
//   4. k_opy<.., true>: inverse FFT along y, keeping rows < Ny;
//   5. k_opx_inv: inverse FFT along x, cropped to the Nx x Ny pixels.
// Padded sizes: powers of two 8..2048 (radix-8 FFTs in LDS) or any length
// up to FFT_RT_MAX (mixed-radix, fft_lds_rt) (fft_core.hpp).
#include <hip/hip_runtime.h>

#include "fft_core.hpp"
#include "ytf_kernels.h"

namespace ytf {

// Round-to-nearest reduction by 2 pi (Julia's rem2pi(x, RoundNearest)) with
// a two-part 2 pi: x - n*C1 is exact for |n| < 2^20, the C2 term restores the
// low bits of 2 pi.
__device__ __forceinline__ double rem2pi_near(double x) {
  const double C1 = 6.283185307179586231995926937088;  // double(2 pi)
  const double C2 = 2.4492935982947063e-16;            // 2 pi - C1
  const double n = rint(x * 0.15915494309189533577);
  double r = __builtin_fma(-n, C1, x);
  r = __builtin_fma(-n, C2, r);
  return r;
}

__device__ __forceinline__ cd finite_or_zero(cd z) { return isfin(z) ? z : cd{0.0, 0.0}; }

// NUFFT points and the node weights times the centring ramp:
//   xf = rem2pi(kx px), yf = rem2pi(ky py), wr = W exp(i (kx x0 + ky y0))
// (_nufft_nodes_2d, mod_ClosedFormField.jl:668-688, with a_x = 0 and
// x0, y0 the offsets of the centre pixel).
__global__ void k_lat_points(int64_t n, const double *__restrict__ kx, const double *__restrict__ ky,
                             const cd *__restrict__ W, double px, double py, double x0, double y0,
                             double *__restrict__ xf, double *__restrict__ yf, cd *__restrict__ wr) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double a = kx[k], b = ky[k];
  xf[k] = rem2pi_near(a * px);
  yf[k] = rem2pi_near(b * py);
  double s, c;
  sincos(a * x0 + b * y0, &s, &c);
  wr[k] = W[k] * cd{c, s};
}

// Strengths of the lateral sum: out[p*tp + e*te][k] = finite_or_zero(w[k] g~[p][e][k]).
// In place (out == g) only with tp = 9, te = 1.
__global__ void k_lat_strengths(int64_t n, int np, const cd *g, const cd *__restrict__ w, cd *out, int64_t tp,
                                int64_t te) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const cd wk = w[k];
  for (int p = 0; p < np; ++p)
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const cd v = g[((int64_t)p * 9 + e) * n + k];
      out[(p * tp + e * te) * n + k] = finite_or_zero(wk * v);
    }
}

// Exact lateral sum.  One 16 x 16 pixel tile and one depth pair per block;
// per chunk of 64 nodes the phases exp(i kx x_i), exp(i ky y_j) of the tile
// and the 9 weighted strengths are staged in LDS, then every thread adds the
// chunk's terms to its pixel's 9 entries.
//   G[i + nx*(j + ny*(p*sp + e*se))]
constexpr int LAT_T = 16, LAT_C = 64;
__global__ __launch_bounds__(256) void k_lat_direct(int64_t n, const double *__restrict__ kx,
                                                    const double *__restrict__ ky, const cd *__restrict__ W,
                                                    const cd *__restrict__ g, const double *__restrict__ xs, int nx,
                                                    const double *__restrict__ ys, int ny, cd *__restrict__ G,
                                                    int64_t sp, int64_t se) {
  __shared__ cd ax[LAT_C][LAT_T], ay[LAT_C][LAT_T], sv[9][LAT_C];
  const int ti = threadIdx.x % LAT_T, tj = threadIdx.x / LAT_T;
  const int i0 = blockIdx.x * LAT_T, j0 = blockIdx.y * LAT_T, p = blockIdx.z;
  const int tid = threadIdx.x;
  cd acc[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) acc[e] = cd{0.0, 0.0};
  const cd *gp = g + (int64_t)p * 9 * n;
  for (int64_t k0 = 0; k0 < n; k0 += LAT_C) {
    const int m = (int)((n - k0) < LAT_C ? (n - k0) : LAT_C);
    for (int e = tid; e < 2 * LAT_C * LAT_T; e += 256) {
      const int h = e / (LAT_C * LAT_T), r = e % (LAT_C * LAT_T);
      const int kk = r / LAT_T, t = r % LAT_T;
      cd v = cd{0.0, 0.0};
      if (kk < m) {
        const int idx = (h ? j0 : i0) + t;
        const bool in = h ? idx < ny : idx < nx;
        if (in) {
          const double ph = h ? ky[k0 + kk] * ys[idx] : kx[k0 + kk] * xs[idx];
          sincos(ph, &v.i, &v.r);
        }
      }
      if (h) ay[kk][t] = v;
      else ax[kk][t] = v;
    }
    for (int e = tid; e < 9 * LAT_C; e += 256) {
      const int ee = e / LAT_C, kk = e % LAT_C;
      sv[ee][kk] = kk < m ? finite_or_zero(W[k0 + kk] * gp[ee * n + k0 + kk]) : cd{0.0, 0.0};
    }
    __syncthreads();
    for (int kk = 0; kk < m; ++kk) {
      const cd ph = ax[kk][ti] * ay[kk][tj];
#pragma unroll
      for (int e = 0; e < 9; ++e) acc[e] = fmac(acc[e], sv[e][kk], ph);
    }
    __syncthreads();
  }
  const int i = i0 + ti, j = j0 + tj;
  if (i < nx && j < ny) {
#pragma unroll
    for (int e = 0; e < 9; ++e) G[i + (int64_t)nx * (j + (int64_t)ny * (p * sp + e * se))] = acc[e];
  }
}

// Per-segment lateral sums at one point per depth pair: the probe values of
// the adaptive k_rho grids (mod_ClosedFormGrid.jl:204-218), i.e. closed_form_G
// of a one-k_rho-node 2D cache, for many k_rho at once.  Nodes come in
// segments of L consecutive nodes (one k_rho, every phi);
//   out[(s np + p) 9 + e] = sum_{j < L} f(W[k] g~[p][e][k]) e^{i kx[k] x[p]} e^{i ky[k] y[p]},
// k = s L + j, f = finite_or_zero: the terms of k_lat_direct for that point.
// One wave per (segment, pair); lanes stride over the segment, then a
// fixed-order shuffle tree adds the lanes.
// The points come in the kernel arguments (up to LatPts::kMax of them), so a
// call needs no upload, and no wait for the kernels still reading an earlier
// upload; more points come from device arrays xp, yp.
__global__ __launch_bounds__(64) void k_lat_segments(int64_t n, int L, const double *__restrict__ kx,
                                                     const double *__restrict__ ky, const cd *__restrict__ W, int np,
                                                     const cd *__restrict__ g, const double *__restrict__ xp,
                                                     const double *__restrict__ yp, LatPts pts, cd *__restrict__ out) {
  const int64_t s = blockIdx.x;
  const int p = blockIdx.y, lane = threadIdx.x;
  const double x = xp ? xp[p] : pts.x[p], y = yp ? yp[p] : pts.y[p];
  const cd *gp = g + (int64_t)p * 9 * n;
  double acc[18];
#pragma unroll
  for (int e = 0; e < 18; ++e) acc[e] = 0.0;
  for (int j = lane; j < L; j += 64) {
    const int64_t k = s * L + j;
    cd ax, ay;
    sincos(kx[k] * x, &ax.i, &ax.r);
    sincos(ky[k] * y, &ay.i, &ay.r);
    const cd ph = ax * ay;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const cd v = fmac(cd{acc[2 * e], acc[2 * e + 1]}, finite_or_zero(W[k] * gp[e * n + k]), ph);
      acc[2 * e] = v.r;
      acc[2 * e + 1] = v.i;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int e = 0; e < 18; ++e) acc[e] += __shfl_xor(acc[e], off);
  if (lane == 0) {
    cd *o = out + (s * np + p) * 9;
#pragma unroll
    for (int e = 0; e < 9; ++e) o[e] = cd{acc[2 * e], acc[2 * e + 1]};
  }
}

// Nodes of one-k_rho-node caches for a batch of k_rho: node k = i L + j
// (k_rho i, phi j) at (kr_i cos phi_j, kr_i sin phi_j) with weight
// W = kr_i w_j / 4 pi^2 x sinc(kx dx/2) sinc(ky dy/2) (the box footprint,
// mod_ClosedFormField.jl:517-524, 723-727; sinc(u) = sin(u)/u, 1 at 0).
__device__ __forceinline__ double sinc_phys(double u) { return u == 0.0 ? 1.0 : sin(u) / u; }

__global__ void k_polar_nodes(int64_t nk, int L, const double *__restrict__ kr, const double *__restrict__ phi,
                              const double *__restrict__ wphi, double dx, double dy, double *__restrict__ kx,
                              double *__restrict__ ky, cd *__restrict__ W) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nk * L) return;
  const int64_t i = k / L;
  const int j = (int)(k - i * L);
  const double r = kr[i], a = r * cos(phi[j]), b = r * sin(phi[j]);
  kx[k] = a;
  ky[k] = b;
  W[k] = cd{r * wphi[j] / (4.0 * 9.8696044010893586188) * (sinc_phys(a * dx / 2) * sinc_phys(b * dy / 2)), 0.0};
}

hipError_t launch_polar_nodes(int64_t nk, int L, const double *kr, const double *phi, const double *wphi, double dx,
                              double dy, double *kx, double *ky, cd *W, hipStream_t s) {
  if (nk <= 0 || L <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_polar_nodes, dim3((unsigned)((nk * L + 255) / 256)), dim3(256), 0, s, nk, L, kr, phi, wphi,
                     dx, dy, kx, ky, W);
  return hipGetLastError();
}

hipError_t launch_lat_segments(int64_t n_seg, int L, const double *kx, const double *ky, const cd *W, int np,
                               const cd *g, const double *xp, const double *yp, const LatPts &pts, cd *out,
                               hipStream_t s) {
  if (n_seg <= 0 || np <= 0) return hipSuccess;
  if (n_seg > 0x7fffffff || np > 65535 || (!xp && np > LatPts::kMax)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_lat_segments, dim3((unsigned)n_seg, (unsigned)np), dim3(64), 0, s, n_seg * L, L, kx, ky, W,
                     np, g, xp, yp, pts, out);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// Padded 2D FFT passes
// --------------------------------------------------------------------------

// Rows per block of the x passes: at most 512 threads and 8 rows (the rows
// give the length of the contiguous segments of the transposed accesses).
template <int PX> struct OpX {
  static constexpr int Q = PX ? 4096 / PX : 1;  // PX = 0: runtime length (shape from rt_shape)
  static constexpr int YT = Q < 4 ? (Q < 1 ? 1 : Q) : (Q > 8 ? 8 : Q);
};

// Forward FFT along x of plane t = blockIdx.y, input rows zero-padded from
// nx_in to PX and rows >= ny_rows skipped; output transposed: out[t][x][y]
// (row stride py).  Input plane t starts at in + (t / pdiv) * pa + (t % pdiv) * pb.
// PX = 0: runtime length rc.n (rc.tpr threads per row, rc.rows rows).
template <int PX>
__global__ __launch_bounds__(1024) void k_opx_fwd(const cd *__restrict__ in, int pdiv, int64_t pa, int64_t pb,
                                                  int in_row, int nx_in, int ny_rows, cd *__restrict__ out, int py,
                                                  int64_t out_plane, const cd *__restrict__ tw, RowCfg rc) {
  const int px = PX ? PX : rc.n;
  const int YT = PX ? OpX<PX>::YT : rc.rows;
  const int TPR = PX ? PX / 8 : rc.tpr;
  const int S = PX ? LdsRow<PX>::S : lds_stride(rc.n);
  extern __shared__ cd lds[];
  const int t = blockIdx.y, tid = threadIdx.x;
  const cd *src = in + (int64_t)(t / pdiv) * pa + (int64_t)(t % pdiv) * pb;
  cd *dst = out + (int64_t)t * out_plane;
  const int y0 = blockIdx.x * YT;
  for (int e = tid; e < YT * px; e += YT * TPR) {
    const int r = e / px, x = e % px, y = y0 + r;
    lds[r * S + lpad(x)] = (x < nx_in && y < ny_rows) ? src[x + (int64_t)in_row * y] : cd{0.0, 0.0};
  }
  __syncthreads();
  row_fft<PX, false>(lds + (tid / TPR) * S, lds + (YT + tid / TPR) * S, tid % TPR, TPR, px, tw);
  for (int e = tid; e < YT * px; e += YT * TPR) {
    const int r = e % YT, x = e / YT, y = y0 + r;
    if (y < ny_rows) dst[y + (int64_t)py * x] = lds[r * S + lpad(x)];
  }
}

// Inverse FFT along x of plane t (in[t][x][y], rows < ny_rows), cropped to
// x < nx_out and stored to out + (t / pdiv) * pa + (t % pdiv) * pb, row stride out_row.
template <int PX>
__global__ __launch_bounds__(1024) void k_opx_inv(const cd *__restrict__ in, int py, int64_t in_plane,
                                                  cd *__restrict__ out, int pdiv, int64_t pa, int64_t pb, int out_row,
                                                  int nx_out, int ny_rows, const cd *__restrict__ tw, RowCfg rc) {
  const int px = PX ? PX : rc.n;
  const int YT = PX ? OpX<PX>::YT : rc.rows;
  const int TPR = PX ? PX / 8 : rc.tpr;
  const int S = PX ? LdsRow<PX>::S : lds_stride(rc.n);
  extern __shared__ cd lds[];
  const int t = blockIdx.y, tid = threadIdx.x;
  const cd *src = in + (int64_t)t * in_plane;
  cd *dst = out + (int64_t)(t / pdiv) * pa + (int64_t)(t % pdiv) * pb;
  const int y0 = blockIdx.x * YT;
  for (int e = tid; e < YT * px; e += YT * TPR) {
    const int r = e % YT, x = e / YT, y = y0 + r;
    lds[r * S + lpad(x)] = y < ny_rows ? src[y + (int64_t)py * x] : cd{0.0, 0.0};
  }
  __syncthreads();
  row_fft<PX, true>(lds + (tid / TPR) * S, lds + (YT + tid / TPR) * S, tid % TPR, TPR, px, tw);
  for (int e = tid; e < YT * px; e += YT * TPR) {
    const int r = e / px, x = e % px, y = y0 + r;
    if (x < nx_out && y < ny_rows) dst[x + (int64_t)out_row * y] = lds[r * S + lpad(x)];
  }
}

// Columns per block of the y pass: PY/8 threads per column, <= 256 threads.
template <int PY> struct OpY {
  static constexpr int TPC = PY ? PY / 8 : 1;  // PY = 0: runtime length (shape from rt_shape)
  static constexpr int XT = (256 / TPC) < 1 ? 1 : ((256 / TPC) > 64 ? 64 : (256 / TPC));
};

// FFT along y, in place on work[t][x][y] (y stride 1, x stride PY): input
// rows >= ny_in are zeros, rows >= ny_out are not stored; the result is
// multiplied by scale, and by (-1)^(x+y) when `sign`.
template <int PY, bool INV>
__global__ __launch_bounds__(256) void k_opy(cd *__restrict__ work, int px, int64_t plane, int ny_in, int ny_out,
                                             double scale, int sign, const cd *__restrict__ tw, RowCfg rc) {
  const int py = PY ? PY : rc.n;
  const int TPC = PY ? OpY<PY>::TPC : rc.tpr;
  const int XT = PY ? OpY<PY>::XT : rc.rows;
  const int S = PY ? LdsRow<PY>::S : lds_stride(rc.n);
  extern __shared__ cd lds[];
  const int t = blockIdx.y, tid = threadIdx.x;
  cd *base = work + (int64_t)t * plane;
  const int x0 = blockIdx.x * XT;
  for (int e = tid; e < XT * py; e += XT * TPC) {
    const int col = e / py, y = e % py, x = x0 + col;
    lds[col * S + lpad(y)] = (x < px && y < ny_in) ? base[y + (int64_t)py * x] : cd{0.0, 0.0};
  }
  __syncthreads();
  row_fft<PY, INV>(lds + (tid / TPC) * S, lds + (XT + tid / TPC) * S, tid % TPC, TPC, py, tw);
  for (int e = tid; e < XT * py; e += XT * TPC) {
    const int col = e / py, y = e % py, x = x0 + col;
    if (x < px && y < ny_out) {
      const double s = (sign && ((x + y) & 1)) ? -scale : scale;
      base[y + (int64_t)py * x] = rmul(s, lds[col * S + lpad(y)]);
    }
  }
}

// --------------------------------------------------------------------------
// z coupling per spectral node
// --------------------------------------------------------------------------

// E[j][r] (+)= sum_i sum_c K[slot(i, j)][r + 3c] J[i][c] on every node, with
// slot = j - i + ns - 1 (Toeplitz), i + j (Hankel) or i + ns*j (dense pairs);
// with a Toeplitz K, H (if not null) adds the Hankel planes H[i + j] in the
// same pass.  J: [ns][3][nn], E: [no][3][nn], K, H: [slot][9][nn].
// Grid: (nodes / 256, observation planes / JT): thread = one spectral node
// and JT observation planes, so a 128^2 padded grid with 16 planes still
// fills the chip; the JT planes share each J load.
template <int JT>
__global__ __launch_bounds__(256) void k_op_mix(int64_t nn, const cd *__restrict__ J, int ns, cd *__restrict__ E,
                                                int no, const cd *__restrict__ K, const cd *__restrict__ H, int kind,
                                                int accumulate) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nn) return;
  const int j0 = blockIdx.y * JT;
  cd a[JT][3];
#pragma unroll
  for (int jj = 0; jj < JT; ++jj) {
    const int j = j0 + jj;
#pragma unroll
    for (int r = 0; r < 3; ++r)
      a[jj][r] = (accumulate && j < no) ? E[(3 * (int64_t)j + r) * nn + k] : cd{0.0, 0.0};
  }
  for (int i = 0; i < ns; ++i) {
    const cd j0v = J[(3 * (int64_t)i + 0) * nn + k], j1v = J[(3 * (int64_t)i + 1) * nn + k],
             j2v = J[(3 * (int64_t)i + 2) * nn + k];
#pragma unroll
    for (int jj = 0; jj < JT; ++jj) {
      const int j = j0 + jj;
      if (j < no) {
        const int slot = kind == 0 ? j - i + ns - 1 : (kind == 1 ? i + j : i + ns * j);
        const cd *Ks = K + (int64_t)slot * 9 * nn + k;
#pragma unroll
        for (int r = 0; r < 3; ++r)
          a[jj][r] = fmac(fmac(fmac(a[jj][r], Ks[r * nn], j0v), Ks[(r + 3) * nn], j1v), Ks[(r + 6) * nn], j2v);
        if (H) {
          const cd *Hs = H + (int64_t)(i + j) * 9 * nn + k;
#pragma unroll
          for (int r = 0; r < 3; ++r)
            a[jj][r] = fmac(fmac(fmac(a[jj][r], Hs[r * nn], j0v), Hs[(r + 3) * nn], j1v), Hs[(r + 6) * nn], j2v);
        }
      }
    }
  }
#pragma unroll
  for (int jj = 0; jj < JT; ++jj) {
    const int j = j0 + jj;
    if (j < no) {
#pragma unroll
      for (int r = 0; r < 3; ++r) E[(3 * (int64_t)j + r) * nn + k] = a[jj][r];
    }
  }
}

// out[o][c] (+)= sum_i M[i + nin*o] in[i][c] on every node (the skeleton fold
// J Ps and expansion Po Es; M is a small complex matrix, read uniformly).
// Grid: (nodes / 256, nout): thread = one node and one output plane.
__global__ __launch_bounds__(256) void k_op_zcomb(int64_t nn, const cd *__restrict__ in, int nin,
                                                  const cd *__restrict__ M, cd *__restrict__ out, int nout,
                                                  int accumulate) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int o = blockIdx.y;
  if (k >= nn) return;
  cd a[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) a[c] = accumulate ? out[(3 * (int64_t)o + c) * nn + k] : cd{0.0, 0.0};
  for (int i = 0; i < nin; ++i) {
    const cd m = M[i + (int64_t)nin * o];
#pragma unroll
    for (int c = 0; c < 3; ++c) a[c] = fmac(a[c], m, in[(3 * (int64_t)i + c) * nn + k]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) out[(3 * (int64_t)o + c) * nn + k] = a[c];
}

// Single source and observation plane, fused: for XT columns x, FFT along y
// of the three source components in[c][x][y] (rows >= ny_in zero), the 3x3
// product with the stored plane K[e][x][y], inverse FFT along y, rows < ny_out
// stored to out[r][x][y].
template <int PY> struct OpYM {
  // columns per block: one for long columns (three blocks fit a CU's LDS and
  // interleave their load, FFT and store phases), more for short ones
  static constexpr int XT = PY >= 512 ? 1 : (PY >= 128 ? 2 : 8);
};
template <int PY>
__global__ __launch_bounds__(1024) void k_opy_mix1(const cd *__restrict__ in, cd *__restrict__ out, int64_t nn,
                                                   const cd *__restrict__ K, int ny_in, int ny_out,
                                                   const cd *__restrict__ tw) {
  constexpr int XT = OpYM<PY>::XT;
  constexpr int TPC = PY / 8;
  constexpr int S = LdsRow<PY>::S;
  extern __shared__ cd lds[];  // [3 * XT][S]
  const int x0 = blockIdx.x * XT, tid = threadIdx.x;
  const int nthr = 3 * XT * TPC;
  for (int e = tid; e < 3 * XT * PY; e += nthr) {
    const int col = e / PY, y = e % PY;
    const int c = col / XT, xi = col % XT;
    lds[col * S + lpad(y)] = y < ny_in ? in[c * nn + y + (int64_t)PY * (x0 + xi)] : cd{0.0, 0.0};
  }
  __syncthreads();
  fft_lds<PY, false>(lds + (tid / TPC) * S, tid % TPC, tw);
  for (int e = tid; e < XT * PY; e += nthr) {
    const int xi = e / PY, y = e % PY;
    const int64_t node = y + (int64_t)PY * (x0 + xi);
    const int yp = lpad(y);
    const cd j0 = lds[(0 * XT + xi) * S + yp], j1 = lds[(1 * XT + xi) * S + yp], j2 = lds[(2 * XT + xi) * S + yp];
    cd o3[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      cd a = K[(r + 0) * nn + node] * j0;
      a = fmac(a, K[(r + 3) * nn + node], j1);
      o3[r] = fmac(a, K[(r + 6) * nn + node], j2);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) lds[(r * XT + xi) * S + yp] = o3[r];
  }
  __syncthreads();
  fft_lds<PY, true>(lds + (tid / TPC) * S, tid % TPC, tw);
  for (int e = tid; e < 3 * XT * PY; e += nthr) {
    const int col = e / PY, y = e % PY;
    const int c = col / XT, xi = col % XT;
    if (y < ny_out) out[c * nn + y + (int64_t)PY * (x0 + xi)] = lds[col * S + lpad(y)];
  }
}

// dst[e][node] += src[e][node] over n complex values (summing operators).
__global__ void k_op_add(int64_t n, const cd *__restrict__ src, cd *__restrict__ dst) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) dst[k] = dst[k] + src[k];
}

// phi-rotation fill of an azimuthally symmetric stack: g0 [np][9][nk] solved
// at phi = 0, g [np][9][nk*nphi] with node k + nk*q at angle phi_q,
// g(phi) = R g0 R^T, R = [[c, -s, 0], [s, c, 0], [0, 0, 1]]; entry a + 3b of a
// node is row a, column b (both indices are vectors and turn alike).  cs:
// cos phi_q at [q], sin phi_q at [nphi + q].
__global__ void k_rot_fill(int64_t nk, int nphi, int np, const cd *__restrict__ g0, const double *__restrict__ cs,
                           cd *__restrict__ g) {
  const int64_t n = nk * nphi;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * np) return;
  const int64_t p = t / n, node = t - p * n;
  const int64_t q = node / nk, k = node - q * nk;
  const double c = cs[q], s = cs[nphi + q];
  const cd *a = g0 + p * 9 * nk + k;
  cd G[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) G[e] = a[e * nk];
  // columns first: H = G R^T (column b of H = sum_j G[., j] R[b][j])
  cd H[9];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const cd gx = G[r], gy = G[r + 3];
    H[r] = rmul(c, gx) - rmul(s, gy);
    H[r + 3] = rmul(s, gx) + rmul(c, gy);
    H[r + 6] = G[r + 6];
  }
  // then rows: out = R H
  cd *o = g + p * 9 * n + node;
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const cd hx = H[3 * b], hy = H[3 * b + 1];
    o[(3 * b) * n] = rmul(c, hx) - rmul(s, hy);
    o[(3 * b + 1) * n] = rmul(s, hx) + rmul(c, hy);
    o[(3 * b + 2) * n] = H[3 * b + 2];
  }
}

// A stored plane [9][x][y] (node y + py*x) to the reference layout
// (Px, Py, 3, 3) = out[x + px*(y + py*e)].
__global__ void k_op_export(int px, int py, const cd *__restrict__ slot, cd *__restrict__ out) {
  const int64_t nn = (int64_t)px * py;
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= 9 * nn) return;
  const int e = (int)(k / nn);
  const int64_t r = k % nn;
  const int x = (int)(r % px), y = (int)(r / px);
  out[k] = slot[e * nn + y + (int64_t)py * x];
}

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_lat_points(int64_t n, const double *kx, const double *ky, const cd *W, double px, double py,
                             double x0, double y0, double *xf, double *yf, cd *wr, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_lat_points, dim3(nblk(n, 256)), dim3(256), 0, s, n, kx, ky, W, px, py, x0, y0, xf, yf, wr);
  return hipGetLastError();
}
hipError_t launch_lat_strengths(int64_t n, int np, const cd *g, const cd *w, cd *out, int64_t tp, int64_t te,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_lat_strengths, dim3(nblk(n, 256)), dim3(256), 0, s, n, np, g, w, out, tp, te);
  return hipGetLastError();
}
hipError_t launch_lat_direct(int64_t n, const double *kx, const double *ky, const cd *W, int np, const cd *g,
                             const double *xs, int nx, const double *ys, int ny, cd *G, int64_t sp, int64_t se,
                             hipStream_t s) {
  const dim3 grid((nx + LAT_T - 1) / LAT_T, (ny + LAT_T - 1) / LAT_T, np);
  hipLaunchKernelGGL(k_lat_direct, grid, dim3(256), 0, s, n, kx, ky, W, g, xs, nx, ys, ny, G, sp, se);
  return hipGetLastError();
}

template <int PX>
static hipError_t opx(bool inv, const cd *in, int pdiv, int64_t pa, int64_t pb, int row, int nx, int ny_rows,
                      cd *out, int py, int64_t plane, int planes, const cd *tw, hipStream_t s) {
  constexpr int YT = OpX<PX>::YT;
  const dim3 grid((ny_rows + YT - 1) / YT, planes), block(YT * (PX / 8));
  const size_t sh = sizeof(cd) * YT * LdsRow<PX>::S;
  const RowCfg rc{PX, PX / 8, YT};
  if (inv)
    hipLaunchKernelGGL(k_opx_inv<PX>, grid, block, sh, s, in, py, plane, out, pdiv, pa, pb, row, nx, ny_rows, tw, rc);
  else
    hipLaunchKernelGGL(k_opx_fwd<PX>, grid, block, sh, s, in, pdiv, pa, pb, row, nx, ny_rows, out, py, plane, tw, rc);
  return hipGetLastError();
}

hipError_t launch_opx(bool inv, int PX, const cd *in, int pdiv, int64_t pa, int64_t pb, int row, int nx, int ny_rows,
                      cd *out, int py, int64_t plane, int planes, const cd *tw, hipStream_t s) {
  if (planes <= 0 || ny_rows <= 0) return hipSuccess;
  switch (PX) {
#define YTF_OPX(N) \
  case N: return opx<N>(inv, in, pdiv, pa, pb, row, nx, ny_rows, out, py, plane, planes, tw, s);
    YTF_OPX(8) YTF_OPX(16) YTF_OPX(32) YTF_OPX(64) YTF_OPX(128) YTF_OPX(256) YTF_OPX(512) YTF_OPX(1024)
    YTF_OPX(2048)
#undef YTF_OPX
  }
  if (PX < 1 || PX > FFT_RT_MAX) return hipErrorInvalidValue;
  const RowCfg rc = rt_shape(PX, 512, 8, 1);
  const dim3 grid((ny_rows + rc.rows - 1) / rc.rows, planes), block(rc.rows * rc.tpr);
  const size_t sh = rt_lds_bytes(rc, 1);
  if (inv)
    hipLaunchKernelGGL(k_opx_inv<0>, grid, block, sh, s, in, py, plane, out, pdiv, pa, pb, row, nx, ny_rows, tw, rc);
  else
    hipLaunchKernelGGL(k_opx_fwd<0>, grid, block, sh, s, in, pdiv, pa, pb, row, nx, ny_rows, out, py, plane, tw, rc);
  return hipGetLastError();
}

template <int PY>
static hipError_t opy(bool inv, cd *work, int px, int64_t plane, int planes, int ny_in, int ny_out, double scale,
                      int sign, const cd *tw, hipStream_t s) {
  constexpr int XT = OpY<PY>::XT, TPC = OpY<PY>::TPC;
  const dim3 grid((px + XT - 1) / XT, planes), block(XT * TPC);
  const size_t sh = sizeof(cd) * XT * LdsRow<PY>::S;
  const RowCfg rc{PY, TPC, XT};
  if (inv)
    hipLaunchKernelGGL((k_opy<PY, true>), grid, block, sh, s, work, px, plane, ny_in, ny_out, scale, sign, tw, rc);
  else
    hipLaunchKernelGGL((k_opy<PY, false>), grid, block, sh, s, work, px, plane, ny_in, ny_out, scale, sign, tw, rc);
  return hipGetLastError();
}

hipError_t launch_opy(bool inv, int PY, cd *work, int px, int64_t plane, int planes, int ny_in, int ny_out,
                      double scale, int sign, const cd *tw, hipStream_t s) {
  if (planes <= 0) return hipSuccess;
  switch (PY) {
#define YTF_OPY(N) \
  case N: return opy<N>(inv, work, px, plane, planes, ny_in, ny_out, scale, sign, tw, s);
    YTF_OPY(8) YTF_OPY(16) YTF_OPY(32) YTF_OPY(64) YTF_OPY(128) YTF_OPY(256) YTF_OPY(512) YTF_OPY(1024)
    YTF_OPY(2048)
#undef YTF_OPY
  }
  if (PY < 1 || PY > FFT_RT_MAX) return hipErrorInvalidValue;
  const RowCfg rc = rt_shape(PY, 256, 64, 1);
  const dim3 grid((px + rc.rows - 1) / rc.rows, planes), block(rc.rows * rc.tpr);
  const size_t sh = rt_lds_bytes(rc, 1);
  if (inv)
    hipLaunchKernelGGL((k_opy<0, true>), grid, block, sh, s, work, px, plane, ny_in, ny_out, scale, sign, tw, rc);
  else
    hipLaunchKernelGGL((k_opy<0, false>), grid, block, sh, s, work, px, plane, ny_in, ny_out, scale, sign, tw, rc);
  return hipGetLastError();
}

template <int PY>
static hipError_t opy_mix1(const cd *in, cd *out, int px, const cd *K, int ny_in, int ny_out, const cd *tw,
                           hipStream_t s) {
  constexpr int XT = OpYM<PY>::XT;
  if (px % XT) return hipErrorInvalidValue;
  const int64_t nn = (int64_t)px * PY;
  const dim3 grid(px / XT), block(3 * XT * (PY / 8));
  const size_t sh = sizeof(cd) * 3 * XT * LdsRow<PY>::S;
  hipLaunchKernelGGL(k_opy_mix1<PY>, grid, block, sh, s, in, out, nn, K, ny_in, ny_out, tw);
  return hipGetLastError();
}

bool opy_mix1_supported(int PY, int px) {
  switch (PY) {
#define YTF_OPYM(N) \
  case N: return px >= 2 && px % OpYM<N>::XT == 0;
    YTF_OPYM(8) YTF_OPYM(16) YTF_OPYM(32) YTF_OPYM(64) YTF_OPYM(128) YTF_OPYM(256) YTF_OPYM(512) YTF_OPYM(1024)
    YTF_OPYM(2048)
#undef YTF_OPYM
  }
  return false;
}

hipError_t launch_opy_mix1(int PY, const cd *in, cd *out, int px, const cd *K, int ny_in, int ny_out, const cd *tw,
                           hipStream_t s) {
  if (px < 2) return hipErrorInvalidValue;
  switch (PY) {
#define YTF_OPYM(N) \
  case N: return opy_mix1<N>(in, out, px, K, ny_in, ny_out, tw, s);
    YTF_OPYM(8) YTF_OPYM(16) YTF_OPYM(32) YTF_OPYM(64) YTF_OPYM(128) YTF_OPYM(256) YTF_OPYM(512) YTF_OPYM(1024)
    YTF_OPYM(2048)
#undef YTF_OPYM
  }
  return hipErrorInvalidValue;
}

hipError_t launch_op_mix(int64_t nn, const cd *J, int ns, cd *E, int no, const cd *K, const cd *H, int kind,
                         int accumulate, hipStream_t s) {
  if (nn == 0 || no == 0) return hipSuccess;
  if (H && kind != OP_MIX_TOEPLITZ) return hipErrorInvalidValue;
  constexpr int JT = 2;
  hipLaunchKernelGGL(k_op_mix<JT>, dim3(nblk(nn, 256), (no + JT - 1) / JT), dim3(256), 0, s, nn, J, ns, E, no, K, H,
                     kind, accumulate);
  return hipGetLastError();
}
hipError_t launch_op_zcomb(int64_t nn, const cd *in, int nin, const cd *M, cd *out, int nout, int accumulate,
                           hipStream_t s) {
  if (nn == 0 || nout == 0) return hipSuccess;
  hipLaunchKernelGGL(k_op_zcomb, dim3(nblk(nn, 256), nout), dim3(256), 0, s, nn, in, nin, M, out, nout, accumulate);
  return hipGetLastError();
}
hipError_t launch_op_add(int64_t n, const cd *src, cd *dst, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_op_add, dim3(nblk(n, 256)), dim3(256), 0, s, n, src, dst);
  return hipGetLastError();
}
hipError_t launch_rot_fill(int64_t nk, int nphi, int np, const cd *g0, const double *cs, cd *g, hipStream_t s) {
  const int64_t n = nk * nphi * (int64_t)np;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rot_fill, dim3(nblk(n, 256)), dim3(256), 0, s, nk, nphi, np, g0, cs, g);
  return hipGetLastError();
}
hipError_t launch_op_export(int px, int py, const cd *slot, cd *out, hipStream_t s) {
  const int64_t n = 9 * (int64_t)px * py;
  hipLaunchKernelGGL(k_op_export, dim3(nblk(n, 256)), dim3(256), 0, s, px, py, slot, out);
  return hipGetLastError();
}

}  // namespace ytf
