// nufft.hip -- 2D type-1 nonuniform FFT, FP64, on gfx950:
//     f[t][k2][k1] = sum_j c[t][j] exp(+i (k1 x_j + k2 y_j)),
//     k1 in [-N1/2, N1/2), k2 in [-N2/2, N2/2) (N even; -floor(N/2).. for odd N),
// the transform of the 2D lateral sum of closed_form_G (FINUFFT type 1,
// iflag +1, mod_ClosedFormField.jl:758-780), to a tolerance eps.
//
// Method (spread, FFT, deconvolve): with x' = x + pi folded into [0, 2 pi),
// h = 2 pi / nf (nf = fine grid size, a power of two >= 2N) and the
// "exponential of semicircle" kernel phi(z) = exp(beta (sqrt(1 - z^2) - 1))
// of half-width a = w h / 2 (w grid points, beta = 2.30 w),
//     b_m = sum_j c_j phi((m h - x'_j) / a)        (periodically, per dimension),
//     f_k = (-1)^k h F_k / psi(k),  F_k = sum_m b_m exp(+2 pi i k m / nf),
//     psi(k) = a int_{-1}^{1} phi(z) cos(k a z) dz,
// separably in 2D.  The spreading is binned by 32 x 32 fine-grid tiles and
// each tile's sorted points are cut into subproblems of at most NU_SUB
// points, so clustered nodes (the folded k-rho nodes of a lateral sum sit
// mostly near the origin) still spread in many blocks: one block per
// (subproblem, transform) accumulates its points (with a halo of the kernel
// half-width) in LDS with FP64 LDS atomics, then adds the tile to its
// transform's fine grid with FP64 global atomics.  A batch of transforms
// shares the sort and runs in one launch per pass.  Two LDS-staged FFT
// passes (rows, then the kept columns) compute only the N1 x N2 wanted modes
// and apply the deconvolution.
#include <hip/hip_runtime.h>

#include "fft_core.hpp"
#include "ytf_kernels.h"

namespace ytf {

constexpr int NU_TS = 32;    // spreading tile (fine-grid points per side)
constexpr int NU_SUB = 512;  // points per spreading subproblem

__device__ __forceinline__ double fold2pi(double x) {
  // x + pi reduced to [0, 2 pi)
  const double tp = 6.283185307179586476925287;
  double u = x + 3.141592653589793238462643;
  u -= tp * floor(u / tp);
  if (u >= tp) u -= tp;
  if (u < 0.0) u = 0.0;
  return u;
}

// The sort is a counting sort by tile without global atomics.  Binning block
// b takes points [b P, (b + 1) P): it computes their fine-grid coordinates
// (u, v) in [0, nf) and tiles, ranks each point among the block's points of
// its tile (LDS histogram), and writes the histogram as row b of H[nbb][ntiles].
// k_nu_colscan sums H's columns into the tile counts and turns each column
// into the offset of block b's run inside its tile, k_nu_scan turns the
// counts into tile starts, and k_nu_scatter places point j at
// start[tile] + H[b][tile] + rank.
constexpr int NU_BIN_T = 1024;      // threads per binning block
constexpr int NU_BIN_BLOCKS = 256;  // at most this many binning blocks

static int64_t nu_bin_chunk(int64_t M) {  // P: a multiple of NU_BIN_T, >= 8192
  int64_t P = (M + NU_BIN_BLOCKS - 1) / NU_BIN_BLOCKS;
  P = (P + NU_BIN_T - 1) / NU_BIN_T * NU_BIN_T;
  return P < 8192 ? 8192 : P;
}
int nu_bin_blocks(int64_t M) {
  const int64_t P = nu_bin_chunk(M);
  return (int)((M + P - 1) / P);
}

__global__ __launch_bounds__(NU_BIN_T) void k_nu_bin(int64_t M, int64_t P, const double *__restrict__ x,
                                                     const double *__restrict__ y, int nf, int nt,
                                                     double *__restrict__ uv, int *__restrict__ tile,
                                                     int *__restrict__ rank, int *__restrict__ H) {
  extern __shared__ int hist[];
  const int ntiles = nt * nt;
  for (int t = threadIdx.x; t < ntiles; t += blockDim.x) hist[t] = 0;
  __syncthreads();
  const double sc = nf / 6.283185307179586476925287;
  const int64_t j1 = min((int64_t)(blockIdx.x + 1) * P, M);
  for (int64_t j = blockIdx.x * P + threadIdx.x; j < j1; j += blockDim.x) {
    double u = fold2pi(x[j]) * sc, v = fold2pi(y[j]) * sc;
    if (u >= nf) u -= nf;
    if (v >= nf) v -= nf;
    int tx = (int)(u / NU_TS), ty = (int)(v / NU_TS);
    tx = tx < nt ? tx : nt - 1;
    ty = ty < nt ? ty : nt - 1;
    uv[2 * j] = u;
    uv[2 * j + 1] = v;
    const int t = tx + nt * ty;
    tile[j] = t;
    rank[j] = atomicAdd(hist + t, 1);
  }
  __syncthreads();
  int *h = H + (int64_t)blockIdx.x * ntiles;
  for (int t = threadIdx.x; t < ntiles; t += blockDim.x) h[t] = hist[t];
}

// Per tile t (a column of H): count[t] = sum over the nbb binning blocks of
// H[b][t], and H[b][t] <- sum over b' < b of H[b'][t] (where block b's run
// starts inside the tile).  A block takes 64 tiles; its 16 waves split the
// binning blocks into 16 contiguous ranges.
__global__ __launch_bounds__(1024) void k_nu_colscan(int ntiles, int nbb, int *__restrict__ H, int *__restrict__ count) {
  __shared__ int part[16][64];
  const int tl = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int t = blockIdx.x * 64 + tl;
  const int per = (nbb + 15) / 16;
  const int b0 = g * per, b1 = min(b0 + per, nbb);
  int s = 0;
  if (t < ntiles)
    for (int b = b0; b < b1; ++b) s += H[(int64_t)b * ntiles + t];
  part[g][tl] = s;
  __syncthreads();
  int off = 0;
  for (int q = 0; q < g; ++q) off += part[q][tl];
  if (t < ntiles) {
    for (int b = b0; b < b1; ++b) {
      int *h = H + (int64_t)b * ntiles + t;
      const int o = *h;
      *h = off;
      off += o;
    }
    if (g == 15) count[t] = off;
  }
}

// Exclusive scan over ntiles values of one block of 1024 threads: out[t] =
// sum of val(t') for t' < t, out[ntiles] = the total.
template <typename F>
__device__ void block_scan(int ntiles, F val, int *__restrict__ out, int *part) {
  const int tid = threadIdx.x;
  const int per = (ntiles + 1023) / 1024;
  int s = 0;
  for (int i = 0; i < per; ++i) {
    const int t = tid * per + i;
    if (t < ntiles) s += val(t);
  }
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = tid ? part[tid - 1] : 0;
  for (int i = 0; i < per; ++i) {
    const int t = tid * per + i;
    if (t < ntiles) {
      out[t] = run;
      run += val(t);
    }
  }
  if (tid == 1023) out[ntiles] = part[1023];
  __syncthreads();
}

// Tile starts and subproblem starts (substart: tile t owns subproblems
// substart[t] .. substart[t + 1] - 1).
__global__ void k_nu_scan(int ntiles, const int *__restrict__ count, int *__restrict__ start,
                          int *__restrict__ substart) {
  __shared__ int part[1024];
  block_scan(ntiles, [&](int t) { return count[t]; }, start, part);
  block_scan(ntiles, [&](int t) { return (count[t] + NU_SUB - 1) / NU_SUB; }, substart, part);
}

// Subproblem s -> (tile, first point): one thread per tile.
__global__ void k_nu_subs(int ntiles, const int *__restrict__ start, const int *__restrict__ substart,
                          int *__restrict__ sub_tile, int *__restrict__ sub_p0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  for (int s = substart[t], p = start[t]; s < substart[t + 1]; ++s, p += NU_SUB) {
    sub_tile[s] = t;
    sub_p0[s] = p;
  }
}

__global__ void k_nu_scatter(int64_t M, int64_t P, int ntiles, const int *__restrict__ tile,
                             const int *__restrict__ rank, const int *__restrict__ H, const int *__restrict__ start,
                             int *__restrict__ order) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= M) return;
  const int t = tile[j];
  order[start[t] + H[(j / P) * ntiles + t] + rank[j]] = (int)j;
}

__global__ void k_nu_zero(int64_t n, cd *__restrict__ g) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) g[i] = cd{0.0, 0.0};
}

// One block per (subproblem, transform): spread the subproblem's points
// into LDS (its tile + halo), then add the LDS tile to the transform's
// periodic fine grid.  blockIdx.x = subproblem (blocks past the total exit),
// blockIdx.y = transform (strengths c + y * c_stride, grid + y * g_stride).
// The kernel values come from its piecewise-polynomial form (hc: tap i of a
// point at offset t = mu - u + W/2 in [0, 1) is sum_d hc[d W + i] s^d with
// s = 2 t - 1, degree W + 2), W (W + 2) FMAs per dimension instead of W
// square roots and exponentials.
template <int W>
__global__ __launch_bounds__(256) void k_nu_spread(int nf, int nt, const double *__restrict__ hc,
                                                   const int *__restrict__ start, const int *__restrict__ substart,
                                                   const int *__restrict__ sub_tile, const int *__restrict__ sub_p0,
                                                   const int *__restrict__ order, const double *__restrict__ uv,
                                                   const cd *__restrict__ c, int64_t c_stride, cd *__restrict__ grid,
                                                   int64_t g_stride) {
  constexpr int NP = W + 3;  // coefficients per tap
  constexpr int H = (W + 1) / 2;
  constexpr int S = NU_TS + 2 * H;
  __shared__ double tre[S * S], tim[S * S];
  const int sub = blockIdx.x;
  if (sub >= substart[nt * nt]) return;
  const int t = sub_tile[sub];
  const int tx = t % nt, ty = t / nt;
  c += blockIdx.y * c_stride;
  grid += blockIdx.y * g_stride;
  for (int i = threadIdx.x; i < S * S; i += blockDim.x) { tre[i] = 0.0; tim[i] = 0.0; }
  __syncthreads();
  const int p0 = sub_p0[sub], p1 = min(p0 + NU_SUB, start[t + 1]);
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int j = order[p];
    const double u = uv[2 * j], v = uv[2 * j + 1];
    const cd cj = c[j];
    const double mu = ceil(u - 0.5 * W), mv = ceil(v - 0.5 * W);
    const double su = 2.0 * (mu - u) + (W - 1.0), sv = 2.0 * (mv - v) + (W - 1.0);
    double ku[W], kv[W];
#pragma unroll
    for (int i = 0; i < W; ++i) ku[i] = kv[i] = hc[(NP - 1) * W + i];
#pragma unroll
    for (int d = NP - 2; d >= 0; --d)
#pragma unroll
      for (int i = 0; i < W; ++i) {
        ku[i] = __builtin_fma(ku[i], su, hc[d * W + i]);
        kv[i] = __builtin_fma(kv[i], sv, hc[d * W + i]);
      }
    const int ou = (int)mu - tx * NU_TS + H, ov = (int)mv - ty * NU_TS + H;
#pragma unroll
    for (int b = 0; b < W; ++b) {
      const double wr = kv[b] * cj.r, wi = kv[b] * cj.i;
      double *rr = tre + (ov + b) * S + ou, *ri = tim + (ov + b) * S + ou;
#pragma unroll
      for (int a = 0; a < W; ++a) {
        atomicAdd(rr + a, ku[a] * wr);
        atomicAdd(ri + a, ku[a] * wi);
      }
    }
  }
  __syncthreads();
  if (p1 == p0) return;
  for (int i = threadIdx.x; i < S * S; i += blockDim.x) {
    const int a = i % S, b = i / S;
    int gu = tx * NU_TS - H + a, gv = ty * NU_TS - H + b;
    gu += gu < 0 ? nf : 0;
    gu -= gu >= nf ? nf : 0;
    gv += gv < 0 ? nf : 0;
    gv -= gv >= nf ? nf : 0;
    double *gp = (double *)(grid + (int64_t)gv * nf + gu);
    if (tre[i] != 0.0) unsafeAtomicAdd(gp, tre[i]);
    if (tim[i] != 0.0) unsafeAtomicAdd(gp + 1, tim[i]);
  }
}

// Pass 1: exp(+) FFT of each fine-grid row v over u; keep modes k1, stored
// transposed: T1[(k1 + N1/2)][v].
template <int NF>
__global__ __launch_bounds__(1024) void k_nu_rows(const cd *__restrict__ grid, cd *__restrict__ T1, int N1,
                                                  const cd *__restrict__ tw) {
  constexpr int YT = XTile<NF>::YT;
  grid += (int64_t)blockIdx.y * NF * NF;  // transform blockIdx.y
  T1 += (int64_t)blockIdx.y * N1 * NF;
  constexpr int TPR = NF / 8;
  extern __shared__ cd lds[];
  const int v0 = blockIdx.x * YT, tid = threadIdx.x;
  constexpr int S = LdsRow<NF>::S;
  for (int e = tid; e < YT * NF; e += YT * TPR) lds[(e / NF) * S + lpad(e % NF)] = grid[(int64_t)v0 * NF + e];
  __syncthreads();
  fft_lds<NF, true>(lds + (tid / TPR) * S, tid % TPR, tw);
  const int h1 = N1 / 2;
  for (int e = tid; e < YT * N1; e += YT * TPR) {
    const int r = e % YT, kk = e / YT;      // kk = k1 + N1/2
    const int k1 = kk - h1;
    const int idx = k1 < 0 ? k1 + NF : k1;
    T1[(int64_t)kk * NF + v0 + r] = lds[r * S + lpad(idx)];
  }
}

// Pass 2: exp(+) FFT over v of each kept k1 row; keep modes k2, deconvolve,
// f[(k2 + N2/2) * N1 + (k1 + N1/2)].
template <int NF>
__global__ __launch_bounds__(1024) void k_nu_cols(const cd *__restrict__ T1, cd *__restrict__ f, int N1, int N2,
                                                  const double *__restrict__ d1, const double *__restrict__ d2,
                                                  const cd *__restrict__ tw) {
  constexpr int YT = XTile<NF>::YT;
  T1 += (int64_t)blockIdx.y * N1 * NF;  // transform blockIdx.y
  f += (int64_t)blockIdx.y * N1 * N2;
  constexpr int TPR = NF / 8;
  extern __shared__ cd lds[];
  const int r0 = blockIdx.x * YT, tid = threadIdx.x;
  const int nrow = N1 - r0 < YT ? N1 - r0 : YT;
  constexpr int S = LdsRow<NF>::S;
  for (int e = tid; e < YT * NF; e += YT * TPR) {
    const int r = e / NF;
    lds[r * S + lpad(e % NF)] = r < nrow ? T1[(int64_t)r0 * NF + e] : cd{0.0, 0.0};
  }
  __syncthreads();
  fft_lds<NF, true>(lds + (tid / TPR) * S, tid % TPR, tw);
  const int h2 = N2 / 2;
  for (int e = tid; e < YT * N2; e += YT * TPR) {
    const int r = e % YT, kk2 = e / YT;
    if (r >= nrow) continue;
    const int k2 = kk2 - h2;
    const int idx = k2 < 0 ? k2 + NF : k2;
    const double s = d1[r0 + r] * d2[kk2];
    f[(int64_t)kk2 * N1 + r0 + r] = rmul(s, lds[r * S + lpad(idx)]);
  }
}

int nu_width(double eps) {
  int w = (int)ceil(log10(1.0 / eps)) + 1;
  if (w < 2) w = 2;
  if (w > 16) w = 16;
  return w + (w & 1);  // even widths only
}

template <int NF>
static hipError_t run_ffts(int nb, const cd *grid, cd *T1, cd *f, int N1, int N2, const double *d1, const double *d2,
                           const cd *tw, hipStream_t s) {
  constexpr int YT = XTile<NF>::YT;
  const size_t sh = sizeof(cd) * YT * LdsRow<NF>::S;
  hipLaunchKernelGGL(k_nu_rows<NF>, dim3(NF / YT, nb), dim3(YT * NF / 8), sh, s, grid, T1, N1, tw);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_nu_cols<NF>, dim3((N1 + YT - 1) / YT, nb), dim3(YT * NF / 8), sh, s, (const cd *)T1, f, N1,
                     N2, d1, d2, tw);
  return hipGetLastError();
}

int nu_max_subproblems(int64_t M, int nf) {
  const int nt = nf / NU_TS;
  return nt * nt + (int)((M + NU_SUB - 1) / NU_SUB);
}

hipError_t launch_nufft_sort(int64_t M, const double *x, const double *y, int nf, double *uv, int *tile, int *rank,
                             int *H, int *count, int *start, int *substart, int *sub_tile, int *sub_p0, int *order,
                             hipStream_t s) {
  const int nt = nf / NU_TS;
  const int ntiles = nt * nt;
  const int nbb = M > 0 ? nu_bin_blocks(M) : 0;
  const int64_t P = nu_bin_chunk(M);
  if (nbb > 0) {
    hipLaunchKernelGGL(k_nu_bin, dim3(nbb), dim3(NU_BIN_T), sizeof(int) * ntiles, s, M, P, x, y, nf, nt, uv, tile,
                       rank, H);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_nu_colscan, dim3((ntiles + 63) / 64), dim3(1024), 0, s, ntiles, nbb, H, count);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_nu_scan, dim3(1), dim3(1024), 0, s, ntiles, (const int *)count, start, substart);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_nu_subs, dim3((ntiles + 255) / 256), dim3(256), 0, s, ntiles, (const int *)start,
                     (const int *)substart, sub_tile, sub_p0);
  if ((e = hipGetLastError()) != hipSuccess || M == 0) return e;
  hipLaunchKernelGGL(k_nu_scatter, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, M, P, ntiles,
                     (const int *)tile, (const int *)rank, (const int *)H, (const int *)start, order);
  return hipGetLastError();
}

hipError_t launch_nufft_spread(int64_t M, int nb, int nf, int w, const double *hc, const int *start,
                               const int *substart, const int *sub_tile, const int *sub_p0, const int *order,
                               const double *uv, const cd *c, int64_t c_stride, cd *grid, hipStream_t s) {
  const int nt = nf / NU_TS;
  const int64_t ng = (int64_t)nf * nf;
  hipLaunchKernelGGL(k_nu_zero, dim3((unsigned)((nb * ng + 255) / 256)), dim3(256), 0, s, nb * ng, grid);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || M == 0) return e;
  const dim3 g(nu_max_subproblems(M, nf), nb);
#define YTF_SPREAD(WW)                                                                                         \
  case WW:                                                                                                     \
    hipLaunchKernelGGL(k_nu_spread<WW>, g, dim3(256), 0, s, nf, nt, hc, start, substart, sub_tile, sub_p0,   \
                       order, uv, c, c_stride, grid, ng);                                                      \
    break;
  switch (w) {
    YTF_SPREAD(2) YTF_SPREAD(4) YTF_SPREAD(6) YTF_SPREAD(8) YTF_SPREAD(10) YTF_SPREAD(12) YTF_SPREAD(14)
    YTF_SPREAD(16)
    default: return hipErrorInvalidValue;
  }
#undef YTF_SPREAD
  return hipGetLastError();
}

hipError_t launch_nufft_fft(int nb, int nf, const cd *grid, cd *T1, cd *f, int N1, int N2, const double *d1,
                            const double *d2, const cd *tw, hipStream_t s) {
  switch (nf) {
    case 64: return run_ffts<64>(nb, grid, T1, f, N1, N2, d1, d2, tw, s);
    case 128: return run_ffts<128>(nb, grid, T1, f, N1, N2, d1, d2, tw, s);
    case 256: return run_ffts<256>(nb, grid, T1, f, N1, N2, d1, d2, tw, s);
    case 512: return run_ffts<512>(nb, grid, T1, f, N1, N2, d1, d2, tw, s);
    case 1024: return run_ffts<1024>(nb, grid, T1, f, N1, N2, d1, d2, tw, s);
    case 2048: return run_ffts<2048>(nb, grid, T1, f, N1, N2, d1, d2, tw, s);
    case 4096: return run_ffts<4096>(nb, grid, T1, f, N1, N2, d1, d2, tw, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ytf
