// fft_kernels.hip -- LDS-staged FP64 FFTs and the fused spectral convolution
// of the uniform-grid Green operator on gfx950.
//
// Reference behaviour: the spectral transform pair of mod_UniformSource.jl
// (_forward_transform :229-230, _inverse_transform :246-248): for a source
// density J on the nx x ny grid,
//     E = (dkx dky nx ny / (2 pi)^2) * ifft( g~ .* conj(ramp) .* (dx dy ramp .* fft(J)) ),
// and since |ramp| = 1 this is E = scale * bfft2(g~ .* fft2(J)) with the
// whole scalar factor (including ifft's 1/(nx ny)) in `scale`.
//
// Three passes over HBM, each LDS-staged (radix-8 Stockham FFTs, one 1D
// transform per N/8 threads):
//   1. fft along x of J (nx, ny, 3) column-major, written transposed into the
//      work buffer laid out [c][x][y] (y fastest);
//   2. per pair of x columns, all three components: fft along y, the 3x3
//      multiply by g~ (stored y-fastest, node index y + ny*x), inverse fft
//      along y -- fused, in place in the work buffer;
//   3. inverse fft along x, transposed back into E (nx, ny, 3).
// Power-of-two sizes 8..2048 use the radix-8 kernels; any other length
// 1..FFT_RT_MAX the runtime mixed-radix transform (fft_lds_rt).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fft_core.hpp"
#include "ytf_kernels.h"

namespace ytf {


// Pass 1: work[c][x][y] = fft_x(J)[x, y, c].  NX = 0: runtime length
// rc.n with rc.tpr threads per row and rc.rows rows per block (the LDS holds
// the rows and their ping-pong rows).
template <int NX>
__global__ __launch_bounds__(1024) void k_fft_x_fwd(const cd *__restrict__ J, cd *__restrict__ work, int ny,
                                                    const cd *__restrict__ tw, RowCfg rc) {
  const int nx = NX ? NX : rc.n;
  const int YT = NX ? XTile<NX>::YT : rc.rows;
  const int S = NX ? LdsRow<NX>::S : lds_stride(rc.n);
  const int TPR = NX ? NX / 8 : rc.tpr;
  extern __shared__ cd lds[];
  const int c = blockIdx.y;
  const int y0 = blockIdx.x * YT;
  const int tid = threadIdx.x;
  // coalesced load of YT rows (rows past ny, in the last block, are zeros and not stored)
  for (int e = tid; e < YT * nx; e += YT * TPR) {
    const int r = e / nx, x = e % nx;
    lds[r * S + lpad(x)] = (y0 + r < ny) ? J[x + (int64_t)nx * ((y0 + r) + (int64_t)ny * c)] : cd{0.0, 0.0};
  }
  __syncthreads();
  row_fft<NX, false>(lds + (tid / TPR) * S, lds + (YT + tid / TPR) * S, tid % TPR, TPR, nx, tw);
  // transposed store: consecutive threads -> consecutive y
  for (int e = tid; e < YT * nx; e += YT * TPR) {
    const int r = e % YT, x = e / YT;
    if (y0 + r < ny) work[(y0 + r) + (int64_t)ny * (x + (int64_t)nx * c)] = lds[r * S + lpad(x)];
  }
}

// Pass 3: E[x, y, c] = bfft_x(work)[c][x][y]
template <int NX>
__global__ __launch_bounds__(1024) void k_fft_x_inv(const cd *__restrict__ work, cd *__restrict__ E, int ny,
                                                    const cd *__restrict__ tw, RowCfg rc) {
  const int nx = NX ? NX : rc.n;
  const int YT = NX ? XTile<NX>::YT : rc.rows;
  const int S = NX ? LdsRow<NX>::S : lds_stride(rc.n);
  const int TPR = NX ? NX / 8 : rc.tpr;
  extern __shared__ cd lds[];
  const int c = blockIdx.y;
  const int y0 = blockIdx.x * YT;
  const int tid = threadIdx.x;
  for (int e = tid; e < YT * nx; e += YT * TPR) {
    const int r = e % YT, x = e / YT;
    lds[r * S + lpad(x)] = (y0 + r < ny) ? work[(y0 + r) + (int64_t)ny * (x + (int64_t)nx * c)] : cd{0.0, 0.0};
  }
  __syncthreads();
  row_fft<NX, true>(lds + (tid / TPR) * S, lds + (YT + tid / TPR) * S, tid % TPR, TPR, nx, tw);
  for (int e = tid; e < YT * nx; e += YT * TPR) {
    const int r = e / nx, x = e % nx;
    if (y0 + r < ny) E[x + (int64_t)nx * ((y0 + r) + (int64_t)ny * c)] = lds[r * S + lpad(x)];
  }
}

template <int N> struct YTile { static constexpr int XT = (N <= 1024) ? 2 : 1; };

// Pass 2: for XT columns x, all c: fft_y, 3x3 multiply by g~ (times scale),
// bfft_y; in place on work[c][x][y].  g~ is [9][nn] with node y + ny*x.
template <int NY>
__global__ __launch_bounds__(NY && NY <= 512 ? 3 * YTile<NY>::XT * (NY / 8) : 1024) void k_fft_y_mul(cd *__restrict__ work, const cd *__restrict__ g, int nx,
                                                    const cd *__restrict__ tw, double scale, RowCfg rc) {
  const int ny = NY ? NY : rc.n;
  const int XT = NY ? YTile<NY>::XT : 1;
  const int TPR = NY ? NY / 8 : rc.tpr;
  const int S = NY ? LdsRow<NY>::S : lds_stride(rc.n);
  extern __shared__ cd lds[];  // [3*XT][S] (+ as many ping-pong rows for NY = 0)
  const int x0 = blockIdx.x * XT;
  const int tid = threadIdx.x;
  const int nthr = 3 * XT * TPR;
  const int64_t nn = (int64_t)nx * ny;
  // With a compile-time length, this thread's nodes' g~ (the larger part of
  // the pass's reads) is loaded into registers first, so that its latency
  // overlaps the loads of the work rows and the forward transform.
  // (lengths up to 512: longer rows leave too few registers for it)
  constexpr bool PRE = NY != 0 && NY <= 512;
  constexpr int NPT = PRE ? (YTile<NY>::XT * NY + 3 * YTile<NY>::XT * (NY / 8) - 1) / (3 * YTile<NY>::XT * (NY / 8)) : 1;
  cd gpre[NPT][9];
  if constexpr (PRE) {
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
      const int e = tid + u * nthr;
      const int xi = e / ny, y = e % ny;
      if (e < XT * ny && x0 + xi < nx) {
        const int64_t node = y + (int64_t)ny * (x0 + xi);
#pragma unroll
        for (int i = 0; i < 9; ++i) gpre[u][i] = g[i * nn + node];
      }
    }
  }
  for (int e = tid; e < 3 * XT * ny; e += nthr) {
    const int col = e / ny, y = e % ny;
    const int c = col / XT, xi = col % XT;
    lds[col * S + lpad(y)] = x0 + xi < nx ? work[y + (int64_t)ny * ((x0 + xi) + (int64_t)nx * c)] : cd{0.0, 0.0};
  }
  __syncthreads();
  row_fft<NY, false>(lds + (tid / TPR) * S, lds + (3 * XT + tid / TPR) * S, tid % TPR, TPR, ny, tw);
  if constexpr (PRE) {
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
      const int e = tid + u * nthr;
      const int xi = e / ny, y = e % ny;
      if (e >= XT * ny || x0 + xi >= nx) continue;
      const int yp = lpad(y);
      const cd j0 = lds[(0 * XT + xi) * S + yp], j1 = lds[(1 * XT + xi) * S + yp], j2 = lds[(2 * XT + xi) * S + yp];
      cd out[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        cd acc = gpre[u][c + 0] * j0;
        acc = fmac(acc, gpre[u][c + 3], j1);
        acc = fmac(acc, gpre[u][c + 6], j2);
        out[c] = rmul(scale, acc);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) lds[(c * XT + xi) * S + yp] = out[c];
    }
  } else {
    for (int e = tid; e < XT * ny; e += nthr) {
      const int xi = e / ny, y = e % ny;
      if (x0 + xi >= nx) continue;
      const int64_t node = y + (int64_t)ny * (x0 + xi);
      const int yp = lpad(y);
      const cd j0 = lds[(0 * XT + xi) * S + yp], j1 = lds[(1 * XT + xi) * S + yp], j2 = lds[(2 * XT + xi) * S + yp];
      cd out[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        cd acc = g[(c + 0) * nn + node] * j0;
        acc = fmac(acc, g[(c + 3) * nn + node], j1);
        acc = fmac(acc, g[(c + 6) * nn + node], j2);
        out[c] = rmul(scale, acc);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) lds[(c * XT + xi) * S + yp] = out[c];
    }
  }
  __syncthreads();
  row_fft<NY, true>(lds + (tid / TPR) * S, lds + (3 * XT + tid / TPR) * S, tid % TPR, TPR, ny, tw);
  for (int e = tid; e < 3 * XT * ny; e += nthr) {
    const int col = e / ny, y = e % ny;
    const int c = col / XT, xi = col % XT;
    if (x0 + xi < nx) work[y + (int64_t)ny * ((x0 + xi) + (int64_t)nx * c)] = lds[col * S + lpad(y)];
  }
}

template <int N>
static hipError_t launch_x(bool inv, const cd *in, cd *out, int ny, const cd *tw, hipStream_t st) {
  constexpr int YT = XTile<N>::YT;
  const dim3 grid((ny + YT - 1) / YT, 3), block(YT * (N / 8));
  const size_t sh = (size_t)YT * LdsRow<N>::S * sizeof(cd);
  const RowCfg rc{N, N / 8, YT};
  if (inv) hipLaunchKernelGGL(k_fft_x_inv<N>, grid, block, sh, st, in, out, ny, tw, rc);
  else hipLaunchKernelGGL(k_fft_x_fwd<N>, grid, block, sh, st, in, out, ny, tw, rc);
  return hipGetLastError();
}
template <int N>
static hipError_t launch_y(cd *work, const cd *g, int nx, const cd *tw, double scale, hipStream_t st) {
  constexpr int XT = YTile<N>::XT;
  const dim3 grid((nx + XT - 1) / XT), block(3 * XT * (N / 8));
  const size_t sh = (size_t)3 * XT * LdsRow<N>::S * sizeof(cd);
  hipLaunchKernelGGL(k_fft_y_mul<N>, grid, block, sh, st, work, g, nx, tw, scale, RowCfg{N, N / 8, XT});
  return hipGetLastError();
}

// Runtime lengths (fft_lds_rt): threads per row ~ n/8 (at most 128), rows
// per block up to 8 within 512 threads and 64 KiB of LDS (rows + ping-pong rows).
RowCfg rt_shape(int n, int max_threads, int max_rows, int rows_per_unit) {
  RowCfg rc;
  rc.n = n;
  rc.tpr = std::min(128, std::max(1, (n + 7) / 8));
  const int S = n + n / 8 + 1;
  const int by_lds = (int)(65536 / (2 * (size_t)rows_per_unit * S * sizeof(cd)));
  rc.rows = std::max(1, std::min({max_rows, max_threads / (rows_per_unit * rc.tpr), by_lds}));
  return rc;
}
size_t rt_lds_bytes(const RowCfg &rc, int rows_per_unit) {
  return 2 * (size_t)rows_per_unit * rc.rows * (rc.n + rc.n / 8 + 1) * sizeof(cd);
}
bool fft_length_ok(int n) {
  return (n >= 8 && n <= 2048 && (n & (n - 1)) == 0) || (n >= 1 && n <= FFT_RT_MAX);
}

static hipError_t dispatch_x(int n, bool inv, const cd *in, cd *out, int ny, const cd *tw, hipStream_t st) {
  switch (n) {
    case 8: return launch_x<8>(inv, in, out, ny, tw, st);
    case 16: return launch_x<16>(inv, in, out, ny, tw, st);
    case 32: return launch_x<32>(inv, in, out, ny, tw, st);
    case 64: return launch_x<64>(inv, in, out, ny, tw, st);
    case 128: return launch_x<128>(inv, in, out, ny, tw, st);
    case 256: return launch_x<256>(inv, in, out, ny, tw, st);
    case 512: return launch_x<512>(inv, in, out, ny, tw, st);
    case 1024: return launch_x<1024>(inv, in, out, ny, tw, st);
    case 2048: return launch_x<2048>(inv, in, out, ny, tw, st);
  }
  if (n < 1 || n > FFT_RT_MAX) return hipErrorInvalidValue;
  const RowCfg rc = rt_shape(n, 512, 8, 1);
  const dim3 grid((ny + rc.rows - 1) / rc.rows, 3), block(rc.rows * rc.tpr);
  const size_t sh = rt_lds_bytes(rc, 1);
  if (inv) hipLaunchKernelGGL(k_fft_x_inv<0>, grid, block, sh, st, in, out, ny, tw, rc);
  else hipLaunchKernelGGL(k_fft_x_fwd<0>, grid, block, sh, st, in, out, ny, tw, rc);
  return hipGetLastError();
}
static hipError_t dispatch_y(int n, cd *work, const cd *g, int nx, const cd *tw, double scale, hipStream_t st) {
  switch (n) {
    case 8: return launch_y<8>(work, g, nx, tw, scale, st);
    case 16: return launch_y<16>(work, g, nx, tw, scale, st);
    case 32: return launch_y<32>(work, g, nx, tw, scale, st);
    case 64: return launch_y<64>(work, g, nx, tw, scale, st);
    case 128: return launch_y<128>(work, g, nx, tw, scale, st);
    case 256: return launch_y<256>(work, g, nx, tw, scale, st);
    case 512: return launch_y<512>(work, g, nx, tw, scale, st);
    case 1024: return launch_y<1024>(work, g, nx, tw, scale, st);
    case 2048: return launch_y<2048>(work, g, nx, tw, scale, st);
  }
  if (n < 1 || n > FFT_RT_MAX) return hipErrorInvalidValue;
  // one column (three component rows) per block
  RowCfg rc = rt_shape(n, 1024, 1, 3);
  rc.rows = 1;
  const dim3 grid(nx), block(3 * rc.tpr);
  hipLaunchKernelGGL(k_fft_y_mul<0>, grid, block, rt_lds_bytes(rc, 3), st, work, g, nx, tw, scale, rc);
  return hipGetLastError();
}

hipError_t launch_spectral_apply(int nx, int ny, const cd *ghat_xy, const cd *J, cd *E, cd *work, const cd *tw_x,
                                 const cd *tw_y, double scale, hipStream_t stream) {
  hipError_t e = dispatch_x(nx, false, J, work, ny, tw_x, stream);
  if (e != hipSuccess) return e;
  e = dispatch_y(ny, work, ghat_xy, nx, tw_y, scale, stream);
  if (e != hipSuccess) return e;
  return dispatch_x(nx, true, work, E, ny, tw_x, stream);
}

}  // namespace ytf
