// fft_core.hpp -- radix-8 Stockham FP64 FFT of one row held in LDS (device
// code shared by the spectral convolution and the nonuniform FFT).
// tw[j] = exp(-2 pi i j / N); INV selects exp(+...) (unnormalised).
#pragma once
#include <hip/hip_runtime.h>

#include "ytf_device.hpp"

namespace ytf {

template <bool INV>
__device__ __forceinline__ cd mul_mi(cd a) {  // a * (-i) forward, a * (+i) inverse
  return INV ? cd{-a.i, a.r} : cd{a.i, -a.r};
}
template <bool INV>
__device__ __forceinline__ void dft2(cd &a0, cd &a1) {
  const cd t = a0;
  a0 = t + a1;
  a1 = t - a1;
}
template <bool INV>
__device__ __forceinline__ void dft4(cd &a0, cd &a1, cd &a2, cd &a3) {
  const cd t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = mul_mi<INV>(a1 - a3);
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}
template <bool INV>
__device__ __forceinline__ void dft8(cd *a) {
  cd e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  cd o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  const double h = 0.70710678118654752440;
  // w8^k o_k, w8 = exp(-+ 2 pi i / 8)
  const cd w1 = INV ? cd{h, h} : cd{h, -h};
  const cd w3 = INV ? cd{-h, h} : cd{-h, -h};
  o1 = o1 * w1;
  o2 = mul_mi<INV>(o2);
  o3 = o3 * w3;
  a[0] = e0 + o0; a[4] = e0 - o0;
  a[1] = e1 + o1; a[5] = e1 - o1;
  a[2] = e2 + o2; a[6] = e2 - o2;
  a[3] = e3 + o3; a[7] = e3 - o3;
}

// LDS layout of a row: element i at lpad(i) = i + i/8 (one pad slot per 8
// elements, so the stride-8 writes of a radix-8 stage fall on distinct
// banks), rows LdsRow<N>::S elements apart (N + N/8 + 1: consecutive rows
// start on different banks, for the transposed loads and stores).
__device__ __forceinline__ int lpad(int i) { return i + (i >> 3); }
template <int N> struct LdsRow { static constexpr int S = N + N / 8 + 1; };

// One Stockham radix-R stage on a row in LDS: current length n, stride s.
// Thread t in [0, N/8) performs 8/R butterflies.  tw[j] = exp(-2 pi i j/N).
template <int N, int R, bool INV>
__device__ __forceinline__ void stage(cd *row, int t, int n, int s, const cd *tw) {
  constexpr int NB = 8 / R;
  const int m = n / R;
  cd a[8];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int b = t + u * (N / 8);
    const int q = b % s, p = b / s;
#pragma unroll
    for (int r = 0; r < R; ++r) a[u * R + r] = row[lpad(q + s * (p + r * m))];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int b = t + u * (N / 8);
    const int q = b % s, p = b / s;
    cd *x = a + u * R;
    if (R == 8) dft8<INV>(x);
    if (R == 4) dft4<INV>(x[0], x[1], x[2], x[3]);
    if (R == 2) dft2<INV>(x[0], x[1]);
    const int step = N / n;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      cd w = tw[(p * k) * step];
      if (INV) w.i = -w.i;
      row[lpad(q + s * (R * p + k))] = k == 0 ? x[0] : x[k] * w;
    }
  }
  __syncthreads();
}

// Full length-N transform of `row` (in LDS, padded layout) by N/8
// cooperating threads; all threads of the block must call it (it
// synchronises the block).
template <int N, bool INV>
__device__ __forceinline__ void fft_lds(cd *row, int t, const cd *tw) {
  int n = N, s = 1;
#pragma unroll
  for (int st = 0; st < 12; ++st) {
    if (n >= 8) {
      stage<N, 8, INV>(row, t, n, s, tw);
      n /= 8;
      s *= 8;
    }
  }
  if (n == 4) stage<N, 4, INV>(row, t, n, s, tw);
  if (n == 2) stage<N, 2, INV>(row, t, n, s, tw);
}

// Rows per block of the row passes (YT rows x N/8 threads each, <= 1024 threads).
template <int N> struct XTile { static constexpr int YT = N == 0 ? 1 : ((8192 / N) < 8 ? (8192 / N) : 8); };

// ---------------------------------------------------------------------------
// Runtime length (any n >= 1): mixed-radix Stockham, the same stage formula
// as `stage` with the radix chosen per stage (8, 4, 2, 3, 5, 7, then the
// smallest remaining prime factor).  Stage: remaining length len = R m,
// stride s (n = len s); butterfly b = q + s p (q < s, p < m) reads
// x[q + s (p + r m)], r < R, and writes
//     y[q + s (R p + k)] = w_len^(p k) sum_r x_r w_R^(r k),   k < R,
// with every root read from the length-n table (w_len^j = tw[j n/len]).
// Stages ping-pong between `row` and `buf` (both padded rows in LDS); the
// result ends in `row`.  tpr cooperating threads (t < tpr) per row; all
// threads of the block must call it.  Used for the lengths that are not
// powers of two (the grids of mod_UniformSource.jl: Nx = ceil(2L/dx) is any
// integer).

__device__ __forceinline__ int lds_stride(int n) { return n + n / 8 + 1; }

__device__ __forceinline__ int fft_radix(int len) {
  if ((len & 7) == 0) return 8;
  if ((len & 3) == 0) return 4;
  if ((len & 1) == 0) return 2;
  if (len % 3 == 0) return 3;
  if (len % 5 == 0) return 5;
  if (len % 7 == 0) return 7;
  int p = 11;
  while (len % p) p += 2;
  return p;
}

template <bool INV>
__device__ __forceinline__ cd root(const cd *tw, int j) {
  cd w = tw[j];
  if (INV) w.i = -w.i;
  return w;
}

template <bool INV>
__device__ __forceinline__ void fft_lds_rt(cd *row, cd *buf, int t, int tpr, int n, const cd *tw) {
  cd *src = row, *dst = buf;
  int len = n, s = 1;
  while (len > 1) {
    const int R = fft_radix(len);
    const int m = len / R, step = n / len, wr = n / R;
    if (R <= 8) {
      for (int b = t; b < n / R; b += tpr) {
        const int q = b % s, p = b / s;
        cd a[8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (r < R) a[r] = src[lpad(q + s * (p + r * m))];
        if (R == 8) {
          dft8<INV>(a);
        } else if (R == 4) {
          dft4<INV>(a[0], a[1], a[2], a[3]);
        } else if (R == 2) {
          dft2<INV>(a[0], a[1]);
        } else {  // 3, 5, 7: direct sum
          cd y[7];
#pragma unroll
          for (int k = 0; k < 7; ++k) {
            if (k < R) {
              cd acc = a[0];
#pragma unroll
              for (int r = 1; r < 7; ++r)
                if (r < R) acc = fmac(acc, a[r], root<INV>(tw, ((r * k) % R) * wr));
              y[k] = acc;
            }
          }
#pragma unroll
          for (int k = 0; k < 7; ++k)
            if (k < R) a[k] = y[k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          if (k < R) dst[lpad(q + s * (R * p + k))] = k == 0 ? a[0] : a[k] * root<INV>(tw, p * k * step);
        }
      }
    } else {  // a prime factor > 7: one output per work item
      for (int o = t; o < n; o += tpr) {
        const int q = o % s, rest = o / s;
        const int k = rest % R, p = rest / R;
        cd acc{0.0, 0.0};
        for (int r = 0; r < R; ++r) acc = fmac(acc, src[lpad(q + s * (p + r * m))], root<INV>(tw, ((r * k) % R) * wr));
        dst[lpad(o)] = k == 0 ? acc : acc * root<INV>(tw, p * k * step);
      }
    }
    __syncthreads();
    cd *tmp = src;
    src = dst;
    dst = tmp;
    len = m;
    s *= R;
  }
  if (src != row) {
    for (int i = t; i < n; i += tpr) row[lpad(i)] = src[lpad(i)];
    __syncthreads();
  }
}

// Compile-time length N (a power of two >= 8: fft_lds) or N == 0 (runtime
// length n: fft_lds_rt with the ping-pong row `buf`).
template <int N, bool INV>
__device__ __forceinline__ void row_fft(cd *row, cd *buf, int t, int tpr, int n, const cd *tw) {
  if constexpr (N > 0) {
    fft_lds<N, INV>(row, t, tw);
  } else {
    fft_lds_rt<INV>(row, buf, t, tpr, n, tw);
  }
}

// Launch shape of a runtime-length pass: n, threads per row and rows (or
// columns) per block.  Chosen on the host by rt_shape.
struct RowCfg {
  int n, tpr, rows;
};
// Longest runtime length (its rows, ping-pong rows and three-component
// columns fit the LDS).
constexpr int FFT_RT_MAX = 1024;
// rows_per_unit LDS rows per block row (3 for the component columns);
// defined in fft_kernels.hip.
RowCfg rt_shape(int n, int max_threads, int max_rows, int rows_per_unit);
size_t rt_lds_bytes(const RowCfg &rc, int rows_per_unit);

}  // namespace ytf
