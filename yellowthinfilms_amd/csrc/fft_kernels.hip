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
// Power-of-two sizes 8..2048.
#include <hip/hip_runtime.h>

#include "fft_core.hpp"
#include "ytf_kernels.h"

namespace ytf {


// Pass 1: work[c][x][y] = fft_x(J)[x, y, c]
template <int NX>
__global__ __launch_bounds__(1024) void k_fft_x_fwd(const cd *__restrict__ J, cd *__restrict__ work, int ny,
                                                    const cd *__restrict__ tw) {
  constexpr int YT = XTile<NX>::YT;
  constexpr int S = LdsRow<NX>::S;
  extern __shared__ cd lds[];
  const int c = blockIdx.y;
  const int y0 = blockIdx.x * YT;
  const int tid = threadIdx.x;
  constexpr int TPR = NX / 8;
  // coalesced load of YT rows
  for (int e = tid; e < YT * NX; e += YT * TPR) {
    const int r = e / NX, x = e % NX;
    lds[r * S + lpad(x)] = J[x + (int64_t)NX * ((y0 + r) + (int64_t)ny * c)];
  }
  __syncthreads();
  fft_lds<NX, false>(lds + (tid / TPR) * S, tid % TPR, tw);
  // transposed store: consecutive threads -> consecutive y
  for (int e = tid; e < YT * NX; e += YT * TPR) {
    const int r = e % YT, x = e / YT;
    work[(y0 + r) + (int64_t)ny * (x + (int64_t)NX * c)] = lds[r * S + lpad(x)];
  }
}

// Pass 3: E[x, y, c] = bfft_x(work)[c][x][y]
template <int NX>
__global__ __launch_bounds__(1024) void k_fft_x_inv(const cd *__restrict__ work, cd *__restrict__ E, int ny,
                                                    const cd *__restrict__ tw) {
  constexpr int YT = XTile<NX>::YT;
  constexpr int S = LdsRow<NX>::S;
  extern __shared__ cd lds[];
  const int c = blockIdx.y;
  const int y0 = blockIdx.x * YT;
  const int tid = threadIdx.x;
  constexpr int TPR = NX / 8;
  for (int e = tid; e < YT * NX; e += YT * TPR) {
    const int r = e % YT, x = e / YT;
    lds[r * S + lpad(x)] = work[(y0 + r) + (int64_t)ny * (x + (int64_t)NX * c)];
  }
  __syncthreads();
  fft_lds<NX, true>(lds + (tid / TPR) * S, tid % TPR, tw);
  for (int e = tid; e < YT * NX; e += YT * TPR) {
    const int r = e / NX, x = e % NX;
    E[x + (int64_t)NX * ((y0 + r) + (int64_t)ny * c)] = lds[r * S + lpad(x)];
  }
}

template <int N> struct YTile { static constexpr int XT = (N <= 1024) ? 2 : 1; };

// Pass 2: for XT columns x, all c: fft_y, 3x3 multiply by g~ (times scale),
// bfft_y; in place on work[c][x][y].  g~ is [9][nn] with node y + ny*x.
template <int NY>
__global__ __launch_bounds__(1024) void k_fft_y_mul(cd *__restrict__ work, const cd *__restrict__ g, int nx,
                                                    const cd *__restrict__ tw, double scale) {
  constexpr int XT = YTile<NY>::XT;
  constexpr int TPR = NY / 8;
  constexpr int S = LdsRow<NY>::S;
  extern __shared__ cd lds[];  // [3*XT][S]
  const int x0 = blockIdx.x * XT;
  const int tid = threadIdx.x;
  const int nthr = 3 * XT * TPR;
  const int64_t nn = (int64_t)nx * NY;
  for (int e = tid; e < 3 * XT * NY; e += nthr) {
    const int col = e / NY, y = e % NY;
    const int c = col / XT, xi = col % XT;
    lds[col * S + lpad(y)] = work[y + (int64_t)NY * ((x0 + xi) + (int64_t)nx * c)];
  }
  __syncthreads();
  fft_lds<NY, false>(lds + (tid / TPR) * S, tid % TPR, tw);
  for (int e = tid; e < XT * NY; e += nthr) {
    const int xi = e / NY, y = e % NY;
    const int64_t node = y + (int64_t)NY * (x0 + xi);
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
  __syncthreads();
  fft_lds<NY, true>(lds + (tid / TPR) * S, tid % TPR, tw);
  for (int e = tid; e < 3 * XT * NY; e += nthr) {
    const int col = e / NY, y = e % NY;
    const int c = col / XT, xi = col % XT;
    work[y + (int64_t)NY * ((x0 + xi) + (int64_t)nx * c)] = lds[col * S + lpad(y)];
  }
}

template <int N>
static hipError_t launch_x(bool inv, const cd *in, cd *out, int ny, const cd *tw, hipStream_t st) {
  constexpr int YT = XTile<N>::YT;
  const dim3 grid(ny / YT, 3), block(YT * (N / 8));
  const size_t sh = (size_t)YT * LdsRow<N>::S * sizeof(cd);
  if (inv) hipLaunchKernelGGL(k_fft_x_inv<N>, grid, block, sh, st, in, out, ny, tw);
  else hipLaunchKernelGGL(k_fft_x_fwd<N>, grid, block, sh, st, in, out, ny, tw);
  return hipGetLastError();
}
template <int N>
static hipError_t launch_y(cd *work, const cd *g, int nx, const cd *tw, double scale, hipStream_t st) {
  constexpr int XT = YTile<N>::XT;
  const dim3 grid(nx / XT), block(3 * XT * (N / 8));
  const size_t sh = (size_t)3 * XT * LdsRow<N>::S * sizeof(cd);
  hipLaunchKernelGGL(k_fft_y_mul<N>, grid, block, sh, st, work, g, nx, tw, scale);
  return hipGetLastError();
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
  return hipErrorInvalidValue;
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
  return hipErrorInvalidValue;
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
