// synth_mfma.hip -- the radial Bessel contraction as a real x complex GEMM on
// the gfx950 matrix cores, with split-precision compensated products.
//
// Same sum as synth.hip (mod_ClosedFormField.jl:344-390 over the row cache
// :304-319), per Bessel order n:
//     R_n[rho, col] = sum_k J_n(kx_k rho) C_n[k, col],
// with C_n the node coefficient columns of that order (real and imaginary
// parts as separate real columns, all depth pairs side by side):
//     n = 0: a, b, e (3 per pair);  n = 1: xz, yz, zx, zy (4);  n = 2: c, d (2).
//
// Precision: J in [-1, 1] is split J = Jh + Jl into two fp16 numbers; each
// coefficient column is scaled by a power of two 2^-e (|C| 2^-e < 1) and split
// the same way.  The product is Jh Ch + Jh Cl + Jl Ch (the dropped Jl Cl is
// below 2^-22 relative), accumulated in FP32 by v_mfma_f32_32x32x16_f16; the
// scale is undone in FP64 when the orders are combined.  The Bessel values
// come from FP64 evaluations; mode solves and g~ stay FP64.
//
// Layouts (all k-fastest, k padded to a multiple of 64 with zeros):
//   Jtab  [order][hi/lo][Mpad][Kpad] fp16     Mpad = radii padded to 128
//   Ctab  [order][hi/lo][Npad_n][Kpad] fp16   column = (p * nc_n + j) * 2 + re/im
//   Cexp  [order][Npad_n] int                 column scale exponents
//   Out   [order][Mpad][Npad_n] fp32
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ytf_kernels.h"

namespace ytf {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16x __attribute__((ext_vector_type(16)));

// J0 and J1 for the split table, whose fp16 pairs keep ~2^-22 of each value
// (absolute 2^-25 below that): an absolute error near 1e-12 is ample.
//   x < 12:  the power series J0 = sum_m (-y)^m / (m!)^2,
//            J1 = (x/2) sum_m (-y)^m / (m! (m+1)!),  y = x^2/4, 31 terms
//            (cancellation in the sum costs at most ~1e-12 at x = 12);
//   x >= 12: Hankel's asymptotic expansion
//            J_nu(x) = sqrt(2/(pi x)) (P_nu cos chi - Q_nu sin chi),
//            chi = x - (nu/2 + 1/4) pi,
//            P_nu = sum_m (-1)^m a_2m / x^2m,  Q_nu = sum_m (-1)^m a_2m+1 / x^2m+1,
//            a_k = prod_{j <= k} (4 nu^2 - (2j - 1)^2) / (k! 8^k), k < 12
//            (truncation 7e-11 at x = 12, far below the table's 2^-22
//            relative split); cos chi and sin chi come from sin x and cos x
//            (sincos_reduced), so no rounding of x - chi enters.
// Against scipy's j0/j1 on [12, 8000] the asymptotic branch is within 7.4e-11.
constexpr int BS_SER = 31, BS_ASY = 6;
struct BesselCoef {
  double s0[BS_SER], s1[BS_SER];                   // 1/(m!)^2, 1/(m! (m+1)!)
  double p0[BS_ASY], q0[BS_ASY], p1[BS_ASY], q1[BS_ASY];  // (-1)^m a_2m, (-1)^m a_2m+1
  double s2[9];                                    // 1/(m! (m+2)!)
};
constexpr BesselCoef make_bessel_coef() {
  BesselCoef c{};
  double f = 1.0;  // m!
  for (int m = 0; m < BS_SER; ++m) {
    if (m > 0) f *= m;
    c.s0[m] = 1.0 / (f * f);
    c.s1[m] = 1.0 / (f * f * (m + 1));
    if (m < 9) c.s2[m] = 1.0 / (f * f * (m + 1) * (m + 2));
  }
  for (int nu = 0; nu < 2; ++nu) {
    double a = 1.0;  // a_0
    for (int k = 0; k < 2 * BS_ASY; ++k) {
      if (k > 0) a *= (4.0 * nu * nu - (2.0 * k - 1.0) * (2.0 * k - 1.0)) / (8.0 * k);
      const int m = k / 2;
      const double v = (m % 2 ? -a : a);
      if (k % 2 == 0) (nu ? c.p1 : c.p0)[m] = v;
      else (nu ? c.q1 : c.q0)[m] = v;
    }
  }
  return c;
}
static constexpr BesselCoef BSC = make_bessel_coef();

// sin x and cos x for 0 <= x < 2^30: x = n pi/2 + r, |r| <= pi/4, by a
// two-term Cody-Waite reduction (the fma makes x - n P1 exact, and P1 + P2
// is pi/2 to 1e-33), then Taylor polynomials in r to r^15 and r^14 (the
// next terms are below 1e-15 on |r| <= pi/4), and the quadrant n mod 4.  A
// fraction of the device library's sincos, whose reduction serves any double.
__device__ inline void sincos_reduced(double x, double &sn, double &cs) {
  const double n = rint(x * 0.63661977236758134308);
  const double r = fma(-n, 6.123233995736766036e-17, fma(-n, 1.5707963267948965580, x));
  const double r2 = r * r;
  double s = -1.0 / 1307674368000.0, c = -1.0 / 87178291200.0;  // r^15/15!, r^14/14!
  s = fma(s, r2, 1.0 / 6227020800.0);
  c = fma(c, r2, 1.0 / 479001600.0);
  s = fma(s, r2, -1.0 / 39916800.0);
  c = fma(c, r2, -1.0 / 3628800.0);
  s = fma(s, r2, 1.0 / 362880.0);
  c = fma(c, r2, 1.0 / 40320.0);
  s = fma(s, r2, -1.0 / 5040.0);
  c = fma(c, r2, -1.0 / 720.0);
  s = fma(s, r2, 1.0 / 120.0);
  c = fma(c, r2, 1.0 / 24.0);
  s = fma(s, r2, -1.0 / 6.0);
  c = fma(c, r2, -0.5);
  s = fma(s * r2, r, r);
  c = fma(c, r2, 1.0);
  const int q = (int)((long long)n & 3);
  sn = q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
  cs = q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}

// rx = 1/ax (used for ax >= 12).
__device__ inline void bessel01_table(double ax, double rx, double &j0, double &j1) {
  if (ax < 12.0) {
    const double y = -0.25 * ax * ax;
    double a = BSC.s0[BS_SER - 1], b = BSC.s1[BS_SER - 1];
#pragma unroll
    for (int m = BS_SER - 2; m >= 0; --m) {
      a = fma(a, y, BSC.s0[m]);
      b = fma(b, y, BSC.s1[m]);
    }
    j0 = a;
    j1 = 0.5 * ax * b;
  } else {
    const double z = rx * rx;
    double p0 = BSC.p0[BS_ASY - 1], q0 = BSC.q0[BS_ASY - 1], p1 = BSC.p1[BS_ASY - 1], q1 = BSC.q1[BS_ASY - 1];
#pragma unroll
    for (int m = BS_ASY - 2; m >= 0; --m) {
      p0 = fma(p0, z, BSC.p0[m]);
      q0 = fma(q0, z, BSC.q0[m]);
      p1 = fma(p1, z, BSC.p1[m]);
      q1 = fma(q1, z, BSC.q1[m]);
    }
    q0 *= rx;
    q1 *= rx;
    double sn, cs;
    if (ax < 1073741824.0) sincos_reduced(ax, sn, cs);
    else sincos(ax, &sn, &cs);
    // sqrt(2/(pi x)) / sqrt(2): cos(x - pi/4) = (cs + sn)/sqrt2, sin(x - pi/4) = (sn - cs)/sqrt2,
    // cos(x - 3pi/4) = (sn - cs)/sqrt2, sin(x - 3pi/4) = -(sn + cs)/sqrt2
    const double amp = sqrt(0.31830988618379067154 * rx);
    j0 = amp * (p0 * (cs + sn) - q0 * (sn - cs));
    j1 = amp * (p1 * (sn - cs) + q1 * (sn + cs));
  }
}

// J0, J1, J2 at x: J2 = 2 J1(x)/x - J0(x), or its power series
// sum_m (-y)^m y / (m! (m+2)!), y = x^2/4, below |x| = 1/2 where that
// difference cancels.
__device__ inline void bessel012m(double x, double &b0, double &b1, double &b2) {
  const double ax = fabs(x);
  // 1/ax by the hardware reciprocal and two Newton steps (within an ulp or
  // so; ax = 0 gives NaN, unused: below 1/2 J2 takes its series and J0, J1
  // theirs)
  double rx = __builtin_amdgcn_rcp(ax);
  rx = fma(rx, fma(-ax, rx, 1.0), rx);
  rx = fma(rx, fma(-ax, rx, 1.0), rx);
  double t;
  bessel01_table(ax, rx, b0, t);
  b1 = x < 0.0 ? -t : t;
  if (ax < 0.5) {
    const double y = 0.25 * ax * ax;
    double s = BSC.s2[8];
#pragma unroll
    for (int m = 7; m >= 0; --m) s = fma(s, -y, BSC.s2[m]);
    b2 = s * y;
  } else {
    b2 = 2.0 * t * rx - b0;
  }
}

// v ~ h + l in fp16: h is v rounded through f32 to fp16 (hardware
// conversions), l the rest rounded the same way.  Rounding through f32 first
// can leave h up to (1/2 + 2^-14) fp16 ulp from v instead of 1/2, so
// |v - h - l| stays within (1 + 2^-12) of the bound of correctly rounded
// halves (2^-22 |v|, or half the fp16 subnormal spacing).
// The empty asm keeps the f32 step: without it the compiler folds the two
// conversions into one f64 -> fp16 rounding, done in software (~40
// instructions each), which made the table kernel compute-bound.
__device__ inline _Float16 to_h(double v) {
  float f = (float)v;
  __asm__ volatile("" : "+v"(f));
  return (_Float16)f;
}
__device__ inline void split_h(double v, _Float16 &h, _Float16 &l) {
  h = to_h(v);
  l = to_h(v - (double)h);
}

// Bessel table: one thread = one radius x 8 consecutive nodes.
__global__ void k_bessel_table(int64_t n, int nr, int Mpad, int64_t Kpad, const double *__restrict__ kx,
                               const double *__restrict__ rho, _Float16 *__restrict__ J) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t kc = Kpad / 8;
  if (t >= (int64_t)Mpad * kc) return;
  const int r = (int)(t / kc);
  const int64_t k0 = (t % kc) * 8;
  h8 v[3][2];
  const double rh = r < nr ? rho[r] : 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int64_t k = k0 + j;
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    if (r < nr && k < n) bessel012m(kx[k] * rh, b0, b1, b2);
    _Float16 h, l;
    split_h(b0, h, l); v[0][0][j] = h; v[0][1][j] = l;
    split_h(b1, h, l); v[1][0][j] = h; v[1][1][j] = l;
    split_h(b2, h, l); v[2][0][j] = h; v[2][1][j] = l;
  }
  const size_t plane = (size_t)Mpad * Kpad;
#pragma unroll
  for (int o = 0; o < 3; ++o)
#pragma unroll
    for (int s = 0; s < 2; ++s) *(h8 *)(J + (size_t)(2 * o + s) * plane + (size_t)r * Kpad + k0) = v[o][s];
}

struct MfmaCols {
  int nc[3];       // complex columns per pair: 3, 4, 2
  int first[3];    // first SC_ index of the order: 0, 3, 7
  int npad[3];     // padded real columns
  size_t ctab[3];  // element offset of each order's hi plane in Ctab
  size_t cexp[3];  // offset in Cexp
  size_t out[3];   // offset in Out
};

// One block = one real coefficient column: its scale exponent, then the split.
// coef: [p][9][n] complex FP64 (synth.hip layout).
__global__ void k_split_columns(int64_t n, int np, int64_t Kpad, MfmaCols cols, const cd *__restrict__ coef,
                                _Float16 *__restrict__ Ctab, int *__restrict__ Cexp) {
  const int o = blockIdx.y;
  const int col = blockIdx.x;
  if (col >= cols.npad[o]) return;
  const int ncols = 2 * cols.nc[o] * np;
  const int ri = col & 1, jj = (col >> 1) % cols.nc[o], p = (col >> 1) / cols.nc[o];
  const bool live = col < ncols;
  const double *src = live ? (const double *)(coef + ((size_t)p * 9 + cols.first[o] + jj) * n) + ri : nullptr;
  __shared__ double red[256];
  double m = 0.0;
  if (live)
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) m = fmax(m, fabs(src[2 * k]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const double mx = red[0];
  // 2^e > max|v|, so the scaled values lie in (-1, 1)
  const int e = (mx > 0.0 && isfinite(mx)) ? ilogb(mx) + 1 : 0;
  if (threadIdx.x == 0) Cexp[cols.cexp[o] + col] = e;
  _Float16 *hi = Ctab + cols.ctab[o] + (size_t)col * Kpad;
  _Float16 *lo = hi + (size_t)cols.npad[o] * Kpad;
  for (int64_t k = threadIdx.x; k < Kpad; k += blockDim.x) {
    double v = (live && k < n) ? ldexp(src[2 * k], -e) : 0.0;
    if (!isfinite(v)) v = 0.0;
    _Float16 h, l;
    split_h(v, h, l);
    hi[k] = h;
    lo[k] = l;
  }
}

// ---------------------------------------------------------------------------
// GEMM: block tile 128 (radii) x 128 (columns), K step 64; 4 waves in 2 x 2,
// each 64 x 64 = 2 x 2 MFMA tiles of 32 x 32.  Global -> registers prefetch of
// the next K step while the current one is multiplied out of LDS.
//
// Work split (stream-K): the (tile, K step) iterations of all tiles, in tile
// order, are cut into gridDim.x equal contiguous ranges, one per block, with
// two blocks per CU.  The 576 tiles of the 8192 x 1152 product would
// otherwise fill 1.125 rounds of the 512 block slots, the last round on an
// eighth of the GPU.  A range covers at least one tile's K steps, so a tile
// is cut at most once: the piece that starts at K step 0 writes output slot 0,
// the rest slot 1, and split[t] tells the combine whether slot 1 holds a part.
// ---------------------------------------------------------------------------
constexpr int BM = 128, BN = 128, BK = 64, LDK = BK + 8;  // LDS row stride (halfs), padded

struct GemmArgs {
  const _Float16 *J;  // Jtab
  const _Float16 *C;  // Ctab
  float *out;         // Out: slot 0, then slot 1 `slot` floats further
  int *split;         // per tile: 1 if slot 1 holds the tile's later K steps
  int64_t slot;
  int64_t Kpad;
  int Mpad;
  int mtiles;
  int ntiles[3];
  int nt_total;  // ntiles[0] + ntiles[1] + ntiles[2]
  MfmaCols cols;
};

__global__ __launch_bounds__(256, 2) void k_bessel_gemm(GemmArgs A) {
  // XCD-grouped ranges: hardware places block b on XCD b % 8; the ranges of
  // one XCD's blocks are contiguous in tile order (radius tile major), so the
  // blocks that share a radius tile's J slab run on the same L2.
  const int nb = gridDim.x, b = blockIdx.x;
  const int per = nb / 8, rem = nb % 8, x = b % 8;
  const int v = x * per + (x < rem ? x : rem) + b / 8;
  const int64_t K = A.Kpad;
  const int ksteps = (int)(K / BK);
  const int64_t iters = (int64_t)A.mtiles * A.nt_total * ksteps;
  int64_t it = v * iters / nb;
  const int64_t it_end = (v + 1) * iters / nb;
  const size_t jplane = (size_t)A.Mpad * K;

  __shared__ _Float16 sAh[BM * LDK], sAl[BM * LDK], sBh[BN * LDK], sBl[BN * LDK];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 31, lk = (lane >> 5) * 8;
  const int hi = lane >> 5;

  while (it < it_end) {
    const int t = (int)(it / ksteps);
    const int k0 = (int)(it % ksteps);
    const int k1 = (int)std::min<int64_t>(ksteps, k0 + (it_end - it));
    const int mt = t / A.nt_total;
    int u = t % A.nt_total, o = 0;
    while (u >= A.ntiles[o]) { u -= A.ntiles[o]; ++o; }
    const int nt = u;
    const int m0 = mt * BM, n0 = nt * BN;
    const _Float16 *Jh = A.J + (size_t)(2 * o) * jplane + (size_t)m0 * K;
    const _Float16 *Jl = Jh + jplane;
    const int npad = A.cols.npad[o];
    const _Float16 *Ch = A.C + A.cols.ctab[o] + (size_t)n0 * K;
    const _Float16 *Cl = Ch + (size_t)npad * K;

    // global load slots: 1024 chunks of 8 halfs per array and K step, 4 per thread
    h8 ra[2][4], rb[2][4];
    auto gload = [&](int64_t kk) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = tid + 256 * i;
        const int row = c >> 3, kc = (c & 7) * 8;
        ra[0][i] = *(const h8 *)(Jh + (size_t)row * K + kk + kc);
        ra[1][i] = *(const h8 *)(Jl + (size_t)row * K + kk + kc);
        rb[0][i] = *(const h8 *)(Ch + (size_t)row * K + kk + kc);
        rb[1][i] = *(const h8 *)(Cl + (size_t)row * K + kk + kc);
      }
    };
    auto sstore = [&]() {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = tid + 256 * i;
        const int row = c >> 3, kc = (c & 7) * 8;
        *(h8 *)(sAh + row * LDK + kc) = ra[0][i];
        *(h8 *)(sAl + row * LDK + kc) = ra[1][i];
        *(h8 *)(sBh + row * LDK + kc) = rb[0][i];
        *(h8 *)(sBl + row * LDK + kc) = rb[1][i];
      }
    };

    f16x acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

    gload((int64_t)k0 * BK);
    for (int ks = k0; ks < k1; ++ks) {
      __syncthreads();
      sstore();
      __syncthreads();
      if (ks + 1 < k1) gload((int64_t)(ks + 1) * BK);
#pragma unroll
      for (int s = 0; s < BK / 16; ++s) {
        h8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int ar = wm * 64 + i * 32 + lr;
          ah[i] = *(const h8 *)(sAh + ar * LDK + s * 16 + lk);
          al[i] = *(const h8 *)(sAl + ar * LDK + s * 16 + lk);
          const int br = wn * 64 + i * 32 + lr;
          bh[i] = *(const h8 *)(sBh + br * LDK + s * 16 + lk);
          bl[i] = *(const h8 *)(sBl + br * LDK + s * 16 + lk);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
          }
      }
    }
    // D layout of the 32 x 32 tile: lane holds column lane % 32, rows
    // 8 g + 4 (lane / 32) + e for element 4 g + e.
    float *out = A.out + (k0 == 0 ? 0 : A.slot) + A.cols.out[o];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = n0 + wn * 64 + j * 32 + lr;
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = m0 + wm * 64 + i * 32 + 8 * g + 4 * hi + e;
            out[(size_t)row * npad + col] = acc[i][j][4 * g + e];
          }
      }
    // a whole tile clears its flag; the later piece of a cut tile sets it
    if (tid == 0) {
      if (k0 != 0) A.split[t] = 1;
      else if (k1 == ksteps) A.split[t] = 0;
    }
    it += k1 - k0;
  }
}

// Combine the three orders into R[p][c + 3q][r] (FP64), undoing the column scales.
__global__ void k_mfma_combine(int np, int nr, MfmaCols cols, const float *__restrict__ out, int64_t slot,
                               const int *__restrict__ split, const int *__restrict__ Cexp, cd *__restrict__ R) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (r >= nr) return;
  int nt_total = 0;
#pragma unroll
  for (int o = 0; o < 3; ++o) nt_total += cols.npad[o] / BN;
  // column counts per order (J0 3, J1 4, J2 2) as compile-time constants, so
  // that s[] stays in registers
  constexpr int NC[3] = {3, 4, 2}, FIRST[3] = {0, 3, 7};
  cd s[9];
  int toff = 0;
#pragma unroll
  for (int o = 0; o < 3; ++o) {
    const float *row = out + cols.out[o] + (size_t)r * cols.npad[o];
    const int *ex = Cexp + cols.cexp[o];
    const int *sp = split + (r / BM) * nt_total + toff;
#pragma unroll
    for (int jj = 0; jj < NC[o]; ++jj) {
      const int c = (p * NC[o] + jj) * 2;
      // the two K parts of a cut tile are added in FP64 (the pair shares a tile)
      double vr = (double)row[c], vi = (double)row[c + 1];
      if (sp[c / BN]) {
        vr += (double)row[slot + c];
        vi += (double)row[slot + c + 1];
      }
      s[FIRST[o] + jj] = cd{ldexp(vr, ex[c]), ldexp(vi, ex[c + 1])};
    }
    toff += cols.npad[o] / BN;
  }
  // index names as in synth.hip: A B E | XZ YZ ZX ZY | C D
  cd *o9 = R + (size_t)p * 9 * nr + r;
  o9[0 * nr] = s[0] + s[7];
  o9[1 * nr] = s[8] - s[1];
  o9[2 * nr] = s[5];
  o9[3 * nr] = s[1] + s[8];
  o9[4 * nr] = s[0] - s[7];
  o9[5 * nr] = s[6];
  o9[6 * nr] = s[3];
  o9[7 * nr] = s[4];
  o9[8 * nr] = s[2];
}

static MfmaCols make_cols(int np, int64_t Kpad, int Mpad) {
  MfmaCols c;
  const int nc[3] = {3, 4, 2}, first[3] = {0, 3, 7};
  size_t ct = 0, ce = 0, co = 0;
  for (int o = 0; o < 3; ++o) {
    c.nc[o] = nc[o];
    c.first[o] = first[o];
    c.npad[o] = (2 * nc[o] * np + BN - 1) / BN * BN;
    c.ctab[o] = ct;
    c.cexp[o] = ce;
    c.out[o] = co;
    ct += (size_t)2 * c.npad[o] * Kpad;
    ce += (size_t)c.npad[o];
    co += (size_t)Mpad * c.npad[o];
  }
  return c;
}

MfmaPlan mfma_plan(int64_t n, int np, int nr) {
  MfmaPlan P;
  P.Kpad = (n + BK - 1) / BK * BK;
  if (P.Kpad == 0) P.Kpad = BK;
  P.Mpad = (nr + BM - 1) / BM * BM;
  MfmaCols c = make_cols(np, P.Kpad, P.Mpad);
  size_t ncol = 0;
  for (int o = 0; o < 3; ++o) ncol += c.npad[o];
  P.jtab_bytes = sizeof(_Float16) * 6 * (size_t)P.Mpad * P.Kpad;
  P.ctab_bytes = sizeof(_Float16) * 2 * ncol * P.Kpad;
  P.cexp_bytes = sizeof(int) * ncol;
  P.out_slot = (int64_t)P.Mpad * ncol;
  P.out_bytes = sizeof(float) * 2 * (size_t)P.out_slot;  // two slots (a cut tile's two K parts)
  P.split_bytes = sizeof(int) * (size_t)(P.Mpad / BM) * (ncol / BN);
  return P;
}

// The four launches of the matrix-core contraction (the C-ABI times each).
hipError_t launch_bessel_table(int64_t n, int np, int nr, const double *kx, const double *rho, void *jtab,
                               hipStream_t s) {
  const MfmaPlan P = mfma_plan(n, np, nr);
  const int64_t nthr = (int64_t)P.Mpad * (P.Kpad / 8);
  hipLaunchKernelGGL(k_bessel_table, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, s, n, nr, P.Mpad, P.Kpad,
                     kx, rho, (_Float16 *)jtab);
  return hipGetLastError();
}

hipError_t launch_split_columns(int64_t n, int np, int nr, const cd *coef, void *ctab, int *cexp, hipStream_t s) {
  const MfmaPlan P = mfma_plan(n, np, nr);
  const MfmaCols cols = make_cols(np, P.Kpad, P.Mpad);
  int maxpad = cols.npad[0];
  for (int o = 1; o < 3; ++o) maxpad = cols.npad[o] > maxpad ? cols.npad[o] : maxpad;
  hipLaunchKernelGGL(k_split_columns, dim3(maxpad, 3), dim3(256), 0, s, n, np, P.Kpad, cols, coef,
                     (_Float16 *)ctab, cexp);
  return hipGetLastError();
}

hipError_t launch_bessel_gemm(int64_t n, int np, int nr, const void *jtab, const void *ctab, float *out, int *split,
                              int n_cu, hipStream_t s) {
  const MfmaPlan P = mfma_plan(n, np, nr);
  const MfmaCols cols = make_cols(np, P.Kpad, P.Mpad);
  GemmArgs A;
  A.J = (const _Float16 *)jtab;
  A.C = (const _Float16 *)ctab;
  A.out = out;
  A.split = split;
  A.slot = P.out_slot;
  A.Kpad = P.Kpad;
  A.Mpad = P.Mpad;
  A.mtiles = P.Mpad / BM;
  A.nt_total = 0;
  for (int o = 0; o < 3; ++o) {
    A.ntiles[o] = cols.npad[o] / BN;
    A.nt_total += A.ntiles[o];
  }
  A.cols = cols;
  // two blocks per CU, and never more blocks than tiles (a range then covers
  // at least one tile's K steps, so no tile is cut twice)
  const int tiles = A.mtiles * A.nt_total;
  const int nblocks = std::max(1, std::min(tiles, 2 * std::max(n_cu, 1)));
  hipLaunchKernelGGL(k_bessel_gemm, dim3(nblocks), dim3(256), 0, s, A);
  return hipGetLastError();
}

hipError_t launch_mfma_combine(int64_t n, int np, int nr, const float *out, const int *split, const int *cexp, cd *R,
                               hipStream_t s) {
  const MfmaPlan P = mfma_plan(n, np, nr);
  const MfmaCols cols = make_cols(np, P.Kpad, P.Mpad);
  hipLaunchKernelGGL(k_mfma_combine, dim3((nr + 255) / 256, np), dim3(256), 0, s, np, nr, cols, out, P.out_slot, split,
                     cexp, R);
  return hipGetLastError();
}

}  // namespace ytf
