// ytf_oracle.cpp -- CPU FP64 restatement of the layered-media Green's-function
// hot path of HortenseBeaudoin/YellowThinFilms (Julia), used as the ORACLE.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (yellowthinfilms_amd/)
// links, loads or calls this file.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it, and only as the checker / CPU baseline.
//
// Every function cites the reference file:line it restates.  The arithmetic is
// templated on the real type R so that the same code runs in plain double and,
// once per workload, with the op-counting real type CR below, which gives the
// algorithmic FP64 operation count per Green k-node used for the roofline.
//
// How each part is pinned (tests/test_oracle_*.py):
//   * eigen4: numpy.linalg.eig (LAPACK zgeev), same normalisation and order;
//   * modes + S-matrix: comparisons/gila_airy/results/{film,stack}.tsv (R_own),
//     TMM/Airy closed form, Fresnel half-space reflection;
//   * T^o and seeds: the identity T^o.[S+;S-] = Modal_amp / Modal_amp_profile
//     (Gz_ClosedForm_Foundation_Validation.jl:38-76); Modal_amp restated here;
//   * spectral dyadic + 1D Bessel synthesis: free-space dipole closed form
//     (Gz_ClosedForm_Point_Validation.jl:164-167), reciprocity, rotation
//     covariance, 1D = 2D-direct-sum on an azimuthally symmetric stack;
//   * anisotropic layers: bulk-Maxwell null vectors of every Berreman
//     eigenpair, the gyrotropic normal-incidence modes (handedness fixed), the
//     seed's particular solution against a bulk 3x3 Maxwell solve, the layer-1
//     dyadic against the angular-spectrum closed form with TE/TM Airy
//     reflections, the 3-layer semi-analytical model
//     (mod_Dipz_SemiAnalytical.jl) and the reference's logged s/p reflection
//     norms of its hyperbolic stack (tests/closed_forms.py);
//   * the interface solve's pivoted-QR fallback: numpy's pinv on rank-0..4
//     systems and on the exactly singular interfaces of a stack at k_z = 0.
//
// ORACLE_MUTANT (default 0) plants one known mistake; tests/test_oracle_mutants.py
// builds each variant on the CPU and shows that the checks above reject it:
//   1  eps_12 and eps_21 exchanged in the Berreman matrix (eps -> eps^T, g -> -g)
//   2  the eps_tz term of the point-source seed dropped
//   3  eps_t and eps_z exchanged in the uniaxial extraordinary k_z
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <type_traits>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace ora {

// ============================================================================
// Op-counting real (add/sub, mul, div, sqrt, transcendental)
// ============================================================================
struct Ops { uint64_t add = 0, mul = 0, div = 0, sqrt = 0, trans = 0; };
static thread_local Ops g_ops;

struct CR {
  double v;
  CR() : v(0.0) {}
  CR(double x) : v(x) {}
};
inline CR operator+(CR a, CR b) { g_ops.add++; return CR(a.v + b.v); }
inline CR operator-(CR a, CR b) { g_ops.add++; return CR(a.v - b.v); }
inline CR operator*(CR a, CR b) { g_ops.mul++; return CR(a.v * b.v); }
inline CR operator/(CR a, CR b) { g_ops.div++; return CR(a.v / b.v); }
inline CR operator-(CR a) { return CR(-a.v); }

inline double val(double x) { return x; }
inline double val(CR x) { return x.v; }
inline double r_sqrt(double x) { return std::sqrt(x); }
inline CR r_sqrt(CR x) { g_ops.sqrt++; return CR(std::sqrt(x.v)); }
inline double r_exp(double x) { return std::exp(x); }
inline CR r_exp(CR x) { g_ops.trans++; return CR(std::exp(x.v)); }
inline double r_cos(double x) { return std::cos(x); }
inline CR r_cos(CR x) { g_ops.trans++; return CR(std::cos(x.v)); }
inline double r_sin(double x) { return std::sin(x); }
inline CR r_sin(CR x) { g_ops.trans++; return CR(std::sin(x.v)); }
inline double r_atan2(double y, double x) { return std::atan2(y, x); }
inline CR r_atan2(CR y, CR x) { g_ops.trans++; return CR(std::atan2(y.v, x.v)); }
inline double r_hypot(double a, double b) { return std::hypot(a, b); }
inline CR r_hypot(CR a, CR b) { g_ops.sqrt++; g_ops.mul += 2; g_ops.add++; return CR(std::hypot(a.v, b.v)); }
inline double r_abs(double x) { return std::fabs(x); }
inline CR r_abs(CR x) { return CR(std::fabs(x.v)); }
inline double r_copysign(double a, double b) { return std::copysign(a, b); }
inline CR r_copysign(CR a, CR b) { return CR(std::copysign(a.v, b.v)); }

// ============================================================================
// Complex over R.  C99 / Julia semantics where signed zeros matter (csqrt).
// ============================================================================
template <class R> struct cx {
  R re, im;
  cx() : re(0.0), im(0.0) {}
  cx(R r, R i) : re(r), im(i) {}
};
template <class R> inline cx<R> C(double re, double im = 0.0) { return cx<R>(R(re), R(im)); }
template <class R> inline cx<R> Cr(R re) { return cx<R>(re, R(0.0)); }
template <class R> inline cx<R> operator+(cx<R> a, cx<R> b) { return cx<R>(a.re + b.re, a.im + b.im); }
template <class R> inline cx<R> operator-(cx<R> a, cx<R> b) { return cx<R>(a.re - b.re, a.im - b.im); }
template <class R> inline cx<R> operator-(cx<R> a) { return cx<R>(-a.re, -a.im); }
template <class R> inline cx<R> operator*(cx<R> a, cx<R> b) {
  return cx<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class R> inline cx<R> rmul(R s, cx<R> a) { return cx<R>(s * a.re, s * a.im); }
// Smith's algorithm for complex division.
template <class R> inline cx<R> operator/(cx<R> a, cx<R> b) {
  double br = val(b.re), bi = val(b.im);
  if (br == 0.0 && bi == 0.0) {
    // Julia: x/0 for complex gives Inf/NaN components
    return cx<R>(R(val(a.re) / 0.0), R(val(a.im) / 0.0));
  }
  if (std::fabs(br) >= std::fabs(bi)) {
    R r = b.im / b.re, den = b.re + b.im * r;
    return cx<R>((a.re + a.im * r) / den, (a.im - a.re * r) / den);
  } else {
    R r = b.re / b.im, den = b.im + b.re * r;
    return cx<R>((a.re * r + a.im) / den, (a.im * r - a.re) / den);
  }
}
template <class R> inline cx<R> conj(cx<R> a) { return cx<R>(a.re, -a.im); }
template <class R> inline R abs2(cx<R> a) { return a.re * a.re + a.im * a.im; }
template <class R> inline R cabs(cx<R> a) { return r_hypot(a.re, a.im); }
template <class R> inline bool isfin(cx<R> a) { return std::isfinite(val(a.re)) && std::isfinite(val(a.im)); }
template <class R> inline cx<R> cexp(cx<R> z) {
  R e = r_exp(z.re);
  return cx<R>(e * r_cos(z.im), e * r_sin(z.im));
}
template <class R> inline cx<R> cis(R x) { return cx<R>(r_cos(x), r_sin(x)); }
template <class R> inline cx<R> csqrt(cx<R> z) {
  double x = val(z.re), y = val(z.im);
  if (x == 0.0 && y == 0.0) return cx<R>(R(0.0), z.im);
  R t = r_sqrt((r_abs(z.re) + cabs(z)) * R(0.5));
  if (x >= 0.0) return cx<R>(t, z.im / (R(2.0) * t));
  (void)y;
  return cx<R>(r_abs(z.im) / (R(2.0) * t), r_copysign(t, z.im));
}
template <class R> inline cx<R> times_i(cx<R> a) { return cx<R>(-a.im, a.re); }

// ============================================================================
// Small column-major matrices
// ============================================================================
template <class R, int M, int N> struct Mat {
  cx<R> a[M * N];
  cx<R> &operator()(int i, int j) { return a[i + M * j]; }
  const cx<R> &operator()(int i, int j) const { return a[i + M * j]; }
};
template <class R, int M, int N> inline Mat<R, M, N> zeros() { return Mat<R, M, N>(); }
template <class R, int N> inline Mat<R, N, N> eye() {
  Mat<R, N, N> A;
  for (int i = 0; i < N; ++i) A(i, i) = C<R>(1.0);
  return A;
}
// StaticArrays' unrolled product: c_ij = a_i1 b_1j + a_i2 b_2j + ...
template <class R, int M, int K, int N>
inline Mat<R, M, N> mul(const Mat<R, M, K> &A, const Mat<R, K, N> &B) {
  Mat<R, M, N> Cm;
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < M; ++i) {
      cx<R> s = A(i, 0) * B(0, j);
      for (int k = 1; k < K; ++k) s = s + A(i, k) * B(k, j);
      Cm(i, j) = s;
    }
  return Cm;
}
// A' * B (adjoint)
template <class R, int M, int K, int N>
inline Mat<R, M, N> adjmul(const Mat<R, K, M> &A, const Mat<R, K, N> &B) {
  Mat<R, M, N> Cm;
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < M; ++i) {
      cx<R> s = conj(A(0, i)) * B(0, j);
      for (int k = 1; k < K; ++k) s = s + conj(A(k, i)) * B(k, j);
      Cm(i, j) = s;
    }
  return Cm;
}
template <class R, int M, int N> inline Mat<R, M, N> add(const Mat<R, M, N> &A, const Mat<R, M, N> &B) {
  Mat<R, M, N> Cm;
  for (int i = 0; i < M * N; ++i) Cm.a[i] = A.a[i] + B.a[i];
  return Cm;
}
template <class R, int M, int N> inline Mat<R, M, N> sub(const Mat<R, M, N> &A, const Mat<R, M, N> &B) {
  Mat<R, M, N> Cm;
  for (int i = 0; i < M * N; ++i) Cm.a[i] = A.a[i] - B.a[i];
  return Cm;
}
template <class R, int M, int N> inline Mat<R, M, N> neg(const Mat<R, M, N> &A) {
  Mat<R, M, N> Cm;
  for (int i = 0; i < M * N; ++i) Cm.a[i] = -A.a[i];
  return Cm;
}
// Diagonal(d) * A  (row scaling) and A * Diagonal(d)
template <class R, int M, int N> inline Mat<R, M, N> dmul(const cx<R> *d, const Mat<R, M, N> &A) {
  Mat<R, M, N> Cm;
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < M; ++i) Cm(i, j) = d[i] * A(i, j);
  return Cm;
}
template <class R, int M, int N> inline Mat<R, M, N> muld(const Mat<R, M, N> &A, const cx<R> *d) {
  Mat<R, M, N> Cm;
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < M; ++i) Cm(i, j) = A(i, j) * d[j];
  return Cm;
}
template <class R, int M, int N> inline bool allfinite(const Mat<R, M, N> &A) {
  for (int i = 0; i < M * N; ++i)
    if (!isfin(A.a[i])) return false;
  return true;
}
// 2x2 block (bi, bj) of a 4x4
template <class R> inline Mat<R, 2, 2> blk(const Mat<R, 4, 4> &S, int bi, int bj) {
  Mat<R, 2, 2> B;
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < 2; ++i) B(i, j) = S(2 * bi + i, 2 * bj + j);
  return B;
}
template <class R> inline Mat<R, 4, 4> from_blocks(const Mat<R, 2, 2> &A11, const Mat<R, 2, 2> &A12,
                                                  const Mat<R, 2, 2> &A21, const Mat<R, 2, 2> &A22) {
  Mat<R, 4, 4> S;
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < 2; ++i) {
      S(i, j) = A11(i, j);
      S(i, j + 2) = A12(i, j);
      S(i + 2, j) = A21(i, j);
      S(i + 2, j + 2) = A22(i, j);
    }
  return S;
}
// StaticArrays 2x2 solve: det = a11 a22 - a12 a21, Cramer per column.
template <class R, int N> inline Mat<R, 2, N> solve2(const Mat<R, 2, 2> &A, const Mat<R, 2, N> &B) {
  cx<R> d = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
  Mat<R, 2, N> X;
  for (int j = 0; j < N; ++j) {
    X(0, j) = (A(1, 1) * B(0, j) - A(0, 1) * B(1, j)) / d;
    X(1, j) = (A(0, 0) * B(1, j) - A(1, 0) * B(0, j)) / d;
  }
  return X;
}
// 4x4 LU with partial pivoting (largest |a|); returns false on an exactly
// zero pivot (Julia's SingularException).
template <class R, int N> inline bool lu_solve4(Mat<R, 4, 4> A, Mat<R, 4, N> B, Mat<R, 4, N> &X) {
  int perm[4] = {0, 1, 2, 3};
  for (int k = 0; k < 4; ++k) {
    int p = k;
    double amax = val(cabs(A(k, k)));
    for (int i = k + 1; i < 4; ++i) {
      double ai = val(cabs(A(i, k)));
      if (ai > amax) { amax = ai; p = i; }
    }
    if (amax == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < 4; ++j) std::swap(A(k, j), A(p, j));
      for (int j = 0; j < N; ++j) std::swap(B(k, j), B(p, j));
      std::swap(perm[k], perm[p]);
    }
    for (int i = k + 1; i < 4; ++i) {
      cx<R> l = A(i, k) / A(k, k);
      A(i, k) = l;
      for (int j = k + 1; j < 4; ++j) A(i, j) = A(i, j) - l * A(k, j);
      for (int j = 0; j < N; ++j) B(i, j) = B(i, j) - l * B(k, j);
    }
  }
  for (int j = 0; j < N; ++j)
    for (int i = 3; i >= 0; --i) {
      cx<R> s = B(i, j);
      for (int k = i + 1; k < 4; ++k) s = s - A(i, k) * X(k, j);
      X(i, j) = s / A(i, i);
    }
  return true;
}
// Column-pivoted Householder QR least-squares solve: the SingularException
// fallback of mod_ScatteringMatrix.jl:35-44, `qr(L, ColumnNorm()) \ R`.
// Julia's QRPivoted solve (LinearAlgebra's ldiv! for QRPivoted: LAPACK xGEQP3,
// then xTZRZF / xORMRZ on the leading rows) returns the minimum-norm
// least-squares solution of the rank-r truncated system, r being the leading
// block whose condition stays below 1/rcond with rcond = 4 eps.  The rank test
// here is |R_kk| <= 4 eps |R_11|; Julia estimates the leading blocks' extreme
// singular values incrementally (xLAIC1).  On the exactly singular systems
// that reach this path both cut at the same rank.  The minimum-norm solution of
// the full-row-rank r x 4 system R y = c is taken from a Householder QR of R^H:
// R^H = Q2 [T; 0], y = Q2 [T^-H c; 0]; then X = P y.
template <class R, int N> inline void qr_solve4(Mat<R, 4, 4> A, Mat<R, 4, N> B, Mat<R, 4, N> &X) {
  int jp[4] = {0, 1, 2, 3};
  double cn[4];
  for (int j = 0; j < 4; ++j) {
    double s = 0;
    for (int i = 0; i < 4; ++i) s += val(abs2(A(i, j)));
    cn[j] = s;
  }
  int rank = 4;
  double rmax = 0;
  // reflect v = (0..0, v_k..v_3) on the columns [k0, 4) of M (4 x NC)
  auto reflect = [](cx<R> *v, int k, auto &M, int ncols, int k0) {
    double vn = 0;
    for (int i = k; i < 4; ++i) vn += val(abs2(v[i]));
    if (vn == 0) return;
    for (int j = k0; j < ncols; ++j) {
      cx<R> s = C<R>(0.0);
      for (int i = k; i < 4; ++i) s = s + conj(v[i]) * M(i, j);
      s = rmul(R(2.0 / vn), s);
      for (int i = k; i < 4; ++i) M(i, j) = M(i, j) - v[i] * s;
    }
  };
  // Householder vector of column x (rows k..3): H x = alpha e_k,
  // alpha = -|x| x_k / |x_k|
  auto house = [](const cx<R> *x, int k, double nrm, cx<R> *v) {
    cx<R> a0 = x[k];
    double aa = val(cabs(a0));
    cx<R> ph = aa == 0 ? C<R>(1.0) : a0 / Cr<R>(R(aa));
    cx<R> alpha = rmul(R(-nrm), ph);
    for (int i = 0; i < 4; ++i) v[i] = C<R>(0.0);
    v[k] = a0 - alpha;
    for (int i = k + 1; i < 4; ++i) v[i] = x[i];
    return alpha;
  };
  for (int k = 0; k < 4; ++k) {
    int p = k;
    for (int j = k + 1; j < 4; ++j)
      if (cn[j] > cn[p]) p = j;
    if (p != k) {
      for (int i = 0; i < 4; ++i) std::swap(A(i, k), A(i, p));
      std::swap(cn[k], cn[p]);
      std::swap(jp[k], jp[p]);
    }
    double nrm2 = 0;
    for (int i = k; i < 4; ++i) nrm2 += val(abs2(A(i, k)));
    double nrm = std::sqrt(nrm2);
    if (k == 0) rmax = nrm;
    if (nrm <= 2.220446049250313e-16 * std::max(rmax, 1e-300) * 4) { rank = k; break; }
    cx<R> col[4], v[4];
    for (int i = 0; i < 4; ++i) col[i] = A(i, k);
    house(col, k, nrm, v);
    reflect(v, k, A, 4, k);
    reflect(v, k, B, N, 0);
    for (int j = k + 1; j < 4; ++j) cn[j] -= val(abs2(A(k, j)));
  }
  // H = R^H (4 x rank), R upper trapezoidal in the leading rows of A
  Mat<R, 4, 4> H;
  for (int i = 0; i < rank; ++i)
    for (int j = 0; j < 4; ++j) H(j, i) = j >= i ? conj(A(i, j)) : C<R>(0.0);
  cx<R> vs[4][4];
  for (int k = 0; k < rank; ++k) {
    double nrm2 = 0;
    for (int i = k; i < 4; ++i) nrm2 += val(abs2(H(i, k)));
    cx<R> col[4];
    for (int i = 0; i < 4; ++i) col[i] = H(i, k);
    house(col, k, std::sqrt(nrm2), vs[k]);
    reflect(vs[k], k, H, rank, k);
  }
  Mat<R, 4, N> Y;
  for (int j = 0; j < N; ++j) {
    // T^H w = c: T^H lower triangular, (T^H)(i, m) = conj(T(m, i)) = conj(H(m, i))
    for (int i = 0; i < rank; ++i) {
      cx<R> s = B(i, j);
      for (int m = 0; m < i; ++m) s = s - conj(H(m, i)) * Y(m, j);
      Y(i, j) = s / conj(H(i, i));
    }
    for (int i = rank; i < 4; ++i) Y(i, j) = C<R>(0.0);
  }
  for (int k = rank - 1; k >= 0; --k) reflect(vs[k], k, Y, N, 0);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < 4; ++i) X(jp[i], j) = Y(i, j);
}

// ============================================================================
// Layers and modes (mod_ModeSolver.jl)
// ============================================================================
enum { EPS_ISO = 0, EPS_UNIAXIAL = 1, EPS_BIAXIAL = 2, EPS_GENERAL = 3 };
enum { MU_ISO = 0, MU_GENERAL = 1 };

template <class R> struct Layer {
  cx<R> eps[9];  // column-major 3x3: eps(i,j) = eps[i + 3j]
  cx<R> mu[9];
  double d;
  int ek, mk;
  cx<R> e(int i, int j) const { return eps[i + 3 * j]; }
  cx<R> m(int i, int j) const { return mu[i + 3 * j]; }
};

template <class R> struct Modes {
  Mat<R, 4, 2> Vf, Vb, Wf, Wb;
  cx<R> Qf[2], Qb[2];
};

// _berreman_core(eps, mu, omega, k) -- mod_ModeSolver.jl:122-193
template <class R> Mat<R, 4, 4> berreman_core(const Layer<R> &L, R omega, R kx, R ky) {
  typedef cx<R> T;
#if ORACLE_MUTANT == 1
  T e11 = L.e(0, 0), e12 = L.e(1, 0), e21 = L.e(0, 1), e22 = L.e(1, 1);
#else
  T e11 = L.e(0, 0), e12 = L.e(0, 1), e21 = L.e(1, 0), e22 = L.e(1, 1);
#endif
  T e13 = L.e(0, 2), e23 = L.e(1, 2), e31 = L.e(2, 0), e32 = L.e(2, 1), e33 = L.e(2, 2);
  T m11 = L.m(0, 0), m12 = L.m(0, 1), m21 = L.m(1, 0), m22 = L.m(1, 1);
  T m13 = L.m(0, 2), m23 = L.m(1, 2), m31 = L.m(2, 0), m32 = L.m(2, 1), m33 = L.m(2, 2);
  T k1 = Cr(kx), k2 = Cr(ky);
  T om = Cr(omega);
  T one = C<R>(1.0);
  T denE = one / e33, denM = one / m33, iw = one / om;
  T ezt1 = e31, ezt2 = e32, mzt1 = m31, mzt2 = m32;
  T Se11 = e11 - e13 * denE * ezt1, Se12 = e12 - e13 * denE * ezt2;
  T Se21 = e21 - e23 * denE * ezt1, Se22 = e22 - e23 * denE * ezt2;
  T Sm11 = m11 - m13 * denM * mzt1, Sm12 = m12 - m13 * denM * mzt2;
  T Sm21 = m21 - m23 * denM * mzt1, Sm22 = m22 - m23 * denM * mzt2;
  T rE1 = denE * ezt1, rE2 = denE * ezt2, sE1 = e13 * denE, sE2 = e23 * denE;
  T rM1 = denM * mzt1, rM2 = denM * mzt2, sM1 = m13 * denM, sM2 = m23 * denM;
  T t2_1 = -conj(k2), t2_2 = conj(k1);
  T t3_1 = iw * k1, t3_2 = iw * k2;
  T RsM1 = sM2, RsM2 = -sM1, RsE1 = sE2, RsE2 = -sE1;
  T B11_11 = RsM1 * t2_1 - k1 * rE1, B11_12 = RsM1 * t2_2 - k1 * rE2;
  T B11_21 = RsM2 * t2_1 - k2 * rE1, B11_22 = RsM2 * t2_2 - k2 * rE2;
  T t1Sm11 = om * Sm21, t1Sm12 = om * Sm22, t1Sm21 = -om * Sm11, t1Sm22 = -om * Sm12;
  T B12_11 = t1Sm11 - t3_1 * denE * t2_1, B12_12 = t1Sm12 - t3_1 * denE * t2_2;
  T B12_21 = t1Sm21 - t3_2 * denE * t2_1, B12_22 = t1Sm22 - t3_2 * denE * t2_2;
  T t1Se11 = om * Se21, t1Se12 = om * Se22, t1Se21 = -om * Se11, t1Se22 = -om * Se12;
  T B21_11 = t3_1 * denM * t2_1 - t1Se11, B21_12 = t3_1 * denM * t2_2 - t1Se12;
  T B21_21 = t3_2 * denM * t2_1 - t1Se21, B21_22 = t3_2 * denM * t2_2 - t1Se22;
  T B22_11 = RsE1 * t2_1 - k1 * rM1, B22_12 = RsE1 * t2_2 - k1 * rM2;
  T B22_21 = RsE2 * t2_1 - k2 * rM1, B22_22 = RsE2 * t2_2 - k2 * rM2;
  Mat<R, 4, 4> B;
  B(0, 0) = B11_11; B(1, 0) = B11_21; B(2, 0) = B21_11; B(3, 0) = B21_21;
  B(0, 1) = B11_12; B(1, 1) = B11_22; B(2, 1) = B21_12; B(3, 1) = B21_22;
  B(0, 2) = B12_11; B(1, 2) = B12_21; B(2, 2) = B22_11; B(3, 2) = B22_21;
  B(0, 3) = B12_12; B(1, 3) = B12_22; B(2, 3) = B22_12; B(3, 3) = B22_22;
  return B;
}

// ----------------------------------------------------------------------------
// General 4x4 complex eigen-decomposition, the stand-in for Julia's
// LinearAlgebra.eigen (LAPACK geevx with balancing, mod_ModeSolver.jl:209):
// balance (diagonal scaling), Householder Hessenberg reduction, shifted QR to
// Schur form, back-substitution eigenvectors, back-transform; each vector
// normalised to unit 2-norm with its largest component real and positive
// (zgeev's convention); eigenpairs sorted by (Re, Im) (Julia's eigsortby).
// Returns false if QR did not converge.
// ----------------------------------------------------------------------------
template <class R> bool eigen4(Mat<R, 4, 4> A, cx<R> *Q, Mat<R, 4, 4> &V) {
  const int n = 4;
  const double ulp = 2.220446049250313e-16;
  // --- balance (EISPACK balanc, radix 2, no permutation) ---
  double D[4] = {1, 1, 1, 1};
  for (int sweep = 0; sweep < 32; ++sweep) {
    bool conv = true;
    for (int i = 0; i < n; ++i) {
      double c = 0, r = 0;
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        c += std::fabs(val(A(j, i).re)) + std::fabs(val(A(j, i).im));
        r += std::fabs(val(A(i, j).re)) + std::fabs(val(A(i, j).im));
      }
      if (c == 0.0 || r == 0.0) continue;
      double g = r / 2.0, f = 1.0, s = c + r;
      while (c < g) { f *= 2.0; c *= 4.0; }
      g = r * 2.0;
      while (c >= g) { f /= 2.0; c /= 4.0; }
      if ((c + r) / f < 0.95 * s) {
        conv = false;
        D[i] *= f;
        for (int j = 0; j < n; ++j) A(i, j) = rmul(R(1.0 / f), A(i, j));
        for (int j = 0; j < n; ++j) A(j, i) = rmul(R(f), A(j, i));
      }
    }
    if (conv) break;
  }
  // --- Hessenberg (Householder), accumulate Z ---
  Mat<R, 4, 4> Z = eye<R, 4>();
  for (int k = 0; k < n - 2; ++k) {
    double xn2 = 0;
    for (int i = k + 1; i < n; ++i) xn2 += val(abs2(A(i, k)));
    if (xn2 == 0.0) continue;
    double xn = std::sqrt(xn2);
    cx<R> x0 = A(k + 1, k);
    double ax0 = val(cabs(x0));
    cx<R> ph = ax0 == 0.0 ? C<R>(1.0) : x0 / Cr<R>(R(ax0));
    cx<R> alpha = rmul(R(-xn), ph);
    cx<R> v[4];
    for (int i = 0; i < n; ++i) v[i] = C<R>(0.0);
    v[k + 1] = x0 - alpha;
    for (int i = k + 2; i < n; ++i) v[i] = A(i, k);
    double vn = 0;
    for (int i = k + 1; i < n; ++i) vn += val(abs2(v[i]));
    if (vn == 0.0) continue;
    R tau = R(2.0 / vn);
    // A <- (I - tau v v^H) A
    for (int j = 0; j < n; ++j) {
      cx<R> s = C<R>(0.0);
      for (int i = k + 1; i < n; ++i) s = s + conj(v[i]) * A(i, j);
      s = rmul(tau, s);
      for (int i = k + 1; i < n; ++i) A(i, j) = A(i, j) - v[i] * s;
    }
    // A <- A (I - tau v v^H);  Z <- Z (I - tau v v^H)
    for (int i = 0; i < n; ++i) {
      cx<R> s = C<R>(0.0), sz = C<R>(0.0);
      for (int j = k + 1; j < n; ++j) { s = s + A(i, j) * v[j]; sz = sz + Z(i, j) * v[j]; }
      s = rmul(tau, s); sz = rmul(tau, sz);
      for (int j = k + 1; j < n; ++j) { A(i, j) = A(i, j) - s * conj(v[j]); Z(i, j) = Z(i, j) - sz * conj(v[j]); }
    }
    for (int i = k + 2; i < n; ++i) A(i, k) = C<R>(0.0);
  }
  // --- shifted QR on the Hessenberg matrix to upper-triangular Schur form ---
  int ihi = n - 1;
  int total_its = 0;
  while (ihi > 0) {
    int its = 0;
    for (;;) {
      int l = 0;
      for (int k = ihi; k >= 1; --k) {
        double hk = std::fabs(val(A(k, k - 1).re)) + std::fabs(val(A(k, k - 1).im));
        double tst = std::fabs(val(A(k - 1, k - 1).re)) + std::fabs(val(A(k - 1, k - 1).im)) +
                     std::fabs(val(A(k, k).re)) + std::fabs(val(A(k, k).im));
        if (tst == 0.0) {
          for (int j = 0; j < n; ++j) tst += std::fabs(val(A(j, k).re)) + std::fabs(val(A(j, k).im));
        }
        if (hk <= ulp * tst) { A(k, k - 1) = C<R>(0.0); l = k; break; }
      }
      if (l == ihi) { ihi -= 1; break; }
      if (++total_its > 120) return false;
      ++its;
      cx<R> mu;
      if (its % 11 == 10) {
        // exceptional shift
        mu = A(ihi, ihi) + rmul(R(0.75), Cr<R>(cabs(A(ihi, ihi - 1))));
      } else {
        cx<R> a = A(ihi - 1, ihi - 1), b = A(ihi - 1, ihi), c = A(ihi, ihi - 1), d = A(ihi, ihi);
        cx<R> half = C<R>(0.5);
        cx<R> tr2 = (a + d) * half;
        cx<R> dd = (a - d) * half;
        cx<R> disc = csqrt(dd * dd + b * c);
        cx<R> l1 = tr2 + disc, l2 = tr2 - disc;
        mu = val(abs2(l1 - d)) <= val(abs2(l2 - d)) ? l1 : l2;
      }
      // explicit shifted QR step on the active block [l, ihi] (full Schur update)
      for (int k = l; k <= ihi; ++k) A(k, k) = A(k, k) - mu;
      double cs[4];
      cx<R> sn[4];
      for (int k = l; k < ihi; ++k) {
        cx<R> a = A(k, k), b = A(k + 1, k);
        double na = val(cabs(a)), nb = val(cabs(b));
        double nrm = std::hypot(na, nb);
        double c;
        cx<R> s;
        if (nrm == 0.0) { c = 1.0; s = C<R>(0.0); }
        else if (na == 0.0) { c = 0.0; s = C<R>(1.0); }
        else {
          c = na / nrm;
          s = rmul(R(1.0 / (na * nrm)), a * conj(b));
        }
        cs[k] = c; sn[k] = s;
        for (int j = k; j < n; ++j) {
          cx<R> x = A(k, j), y = A(k + 1, j);
          A(k, j) = rmul(R(c), x) + s * y;
          A(k + 1, j) = rmul(R(c), y) - conj(s) * x;
        }
      }
      for (int k = l; k < ihi; ++k) {
        double c = cs[k];
        cx<R> s = sn[k];
        for (int i = 0; i <= std::min(k + 1, ihi); ++i) {
          cx<R> x = A(i, k), y = A(i, k + 1);
          A(i, k) = rmul(R(c), x) + y * conj(s);
          A(i, k + 1) = rmul(R(c), y) - x * s;
        }
        for (int i = 0; i < n; ++i) {
          cx<R> x = Z(i, k), y = Z(i, k + 1);
          Z(i, k) = rmul(R(c), x) + y * conj(s);
          Z(i, k + 1) = rmul(R(c), y) - x * s;
        }
      }
      for (int k = l; k <= ihi; ++k) A(k, k) = A(k, k) + mu;
    }
  }
  // --- eigenvectors of the triangular T, back-transformed ---
  cx<R> lam[4];
  for (int k = 0; k < n; ++k) lam[k] = A(k, k);
  double tnorm = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) tnorm = std::max(tnorm, val(cabs(A(i, j))));
  Mat<R, 4, 4> Vt;
  for (int k = 0; k < n; ++k) {
    cx<R> x[4];
    for (int i = 0; i < n; ++i) x[i] = C<R>(0.0);
    x[k] = C<R>(1.0);
    double smin = std::max(ulp * val(cabs(lam[k])), 1e-290);
    if (smin == 0.0) smin = 1e-290;
    for (int j = k - 1; j >= 0; --j) {
      cx<R> s = C<R>(0.0);
      for (int l2 = j + 1; l2 <= k; ++l2) s = s + A(j, l2) * x[l2];
      cx<R> den = A(j, j) - lam[k];
      if (val(cabs(den)) < smin) den = C<R>(smin);
      x[j] = -(s / den);
    }
    for (int i = 0; i < n; ++i) {
      cx<R> s = C<R>(0.0);
      for (int j = 0; j <= k; ++j) s = s + Z(i, j) * x[j];
      Vt(i, k) = rmul(R(D[i]), s);  // undo balancing
    }
  }
  (void)tnorm;
  // normalise (zgeev): unit 2-norm, largest |component| real positive
  for (int k = 0; k < n; ++k) {
    R s2 = abs2(Vt(0, k));
    for (int i = 1; i < n; ++i) s2 = s2 + abs2(Vt(i, k));
    R inv = R(1.0) / r_sqrt(s2);
    double best = 0;
    for (int i = 0; i < n; ++i) {
      Vt(i, k) = rmul(inv, Vt(i, k));
      best = std::max(best, val(abs2(Vt(i, k))));
    }
    // the largest component (first of any near-tie, so rounding cannot
    // change the choice) is made real and positive
    int im = 0;
    for (int i = 0; i < n; ++i)
      if (val(abs2(Vt(i, k))) >= best * (1.0 - 1e-10)) { im = i; break; }
    cx<R> p = Vt(im, k);
    R ap = cabs(p);
    cx<R> rot = cx<R>(p.re / ap, -(p.im / ap));
    for (int i = 0; i < n; ++i) Vt(i, k) = Vt(i, k) * rot;
    Vt(im, k).im = R(0.0);
  }
  // sort by (Re, Im)
  int ord[4] = {0, 1, 2, 3};
  std::stable_sort(ord, ord + 4, [&](int a, int b) {
    double ar = val(lam[a].re), br = val(lam[b].re);
    if (ar != br) return ar < br;
    return val(lam[a].im) < val(lam[b].im);
  });
  for (int k = 0; k < n; ++k) {
    Q[k] = lam[ord[k]];
    for (int i = 0; i < n; ++i) V(i, k) = Vt(i, ord[k]);
  }
  return true;
}

// _eigen_QV closed form, isotropic/uniaxial eps with isotropic mu
// -- mod_ModeSolver.jl:243-265
template <class R> void eigen_closed(const Layer<R> &L, R omega, R kx, R ky, cx<R> *Q, Mat<R, 4, 4> &V) {
  typedef cx<R> T;
  T et = L.e(0, 0), ez = L.e(2, 2), mr = L.m(0, 0);
  T wc = Cr(omega);
  R kr = r_hypot(kx, ky);
  R phi = val(kr) == 0.0 ? R(0.0) : r_atan2(ky, kx);
  R cp = r_cos(phi), sp = r_sin(phi);
  T kr2 = Cr<R>(kr * kr);
  T kzo = csqrt(wc * wc * et * mr - kr2);
#if ORACLE_MUTANT == 3
  T kze = csqrt(ez * (wc * wc * mr - kr2 / et));
#else
  T kze = csqrt(et * (wc * wc * mr - kr2 / ez));
#endif
  T A = C<R>(1.0) / (wc * mr);
  T Bn = C<R>(1.0) / (wc * et);
  T c = Cr(cp), s = Cr(sp);
  // columns TEf, TEb, TMf, TMb
  V(0, 0) = -s; V(1, 0) = c; V(2, 0) = -(A * kzo * c); V(3, 0) = -(A * kzo * s);
  V(0, 1) = -s; V(1, 1) = c; V(2, 1) = A * kzo * c; V(3, 1) = A * kzo * s;
  V(0, 2) = Bn * kze * c; V(1, 2) = Bn * kze * s; V(2, 2) = -s; V(3, 2) = c;
  V(0, 3) = -(Bn * kze * c); V(1, 3) = -(Bn * kze * s); V(2, 3) = -s; V(3, 3) = c;
  Q[0] = kzo; Q[1] = -kzo; Q[2] = kze; Q[3] = -kze;
}

// _eigen_QV dispatcher -- mod_ModeSolver.jl:203-211, 243
template <class R> bool eigen_QV(const Layer<R> &L, R omega, R kx, R ky, cx<R> *Q, Mat<R, 4, 4> &V) {
  if ((L.ek == EPS_ISO || L.ek == EPS_UNIAXIAL) && L.mk == MU_ISO) {
    eigen_closed(L, omega, kx, ky, Q, V);
    return true;
  }
  return eigen4(berreman_core(L, omega, kx, ky), Q, V);
}

// W = R4s * V with R4s = [0 0 0 1; 0 0 -1 0; 0 -1 0 0; 1 0 0 0] (mod_ModeSolver.jl:17)
template <class R, int N> inline Mat<R, 4, N> R4s_mul(const Mat<R, 4, N> &V) {
  Mat<R, 4, N> W;
  for (int j = 0; j < N; ++j) {
    W(0, j) = V(3, j);
    W(1, j) = -V(2, j);
    W(2, j) = -V(1, j);
    W(3, j) = V(0, j);
  }
  return W;
}

// _modes_layer -- mod_ModeSolver.jl:274-326
template <class R> bool modes_layer(const Layer<R> &L, R omega, R kx, R ky, Modes<R> &M) {
  const double tol = 1.4901161193847656e-08;  // sqrt(eps(Float64))
  cx<R> Qs[4];
  Mat<R, 4, 4> Vs;
  if (!eigen_QV(L, omega, kx, ky, Qs, Vs)) return false;
  Mat<R, 4, 4> Ws = R4s_mul(Vs);
  double Pz[4];
  for (int j = 0; j < 4; ++j) {
    cx<R> dt = conj(Vs(0, j)) * Ws(0, j);
    for (int i = 1; i < 4; ++i) dt = dt + conj(Vs(i, j)) * Ws(i, j);
    Pz[j] = val(R(0.25) * dt.re);
  }
  M = Modes<R>();
  int idf = 0, idb = 0, ndeg = 0, deg[4];
  auto putf = [&](int j) {
    for (int i = 0; i < 4; ++i) { M.Vf(i, idf) = Vs(i, j); M.Wf(i, idf) = Ws(i, j); }
    M.Qf[idf] = Qs[j]; idf++;
  };
  auto putb = [&](int j) {
    for (int i = 0; i < 4; ++i) { M.Vb(i, idb) = Vs(i, j); M.Wb(i, idb) = Ws(i, j); }
    M.Qb[idb] = Qs[j]; idb++;
  };
  for (int j = 0; j < 4; ++j) {
    double iq = val(Qs[j].im);
    if (std::fabs(Pz[j]) > tol) {
      if (Pz[j] > 0) { if (idf < 2) putf(j); else return false; }
      else { if (idb < 2) putb(j); else return false; }
    } else if (std::fabs(iq) > tol) {
      if (iq > tol) { if (idf < 2) putf(j); else return false; }
      else if (iq < -tol) { if (idb < 2) putb(j); else return false; }
    } else {
      deg[ndeg++] = j;
    }
  }
  for (int i = 0; i < ndeg; ++i) {
    int j = deg[i];
    if (idf < 2) putf(j);
    else if (idb < 2) putb(j);
  }
  return true;
}

// _match_modes (eigpenalty = 0) -- mod_ModeSolver.jl:333-349
template <class R> void match_modes(const Mat<R, 4, 2> &Wa, Mat<R, 4, 2> &Vb, Mat<R, 4, 2> &Wb, cx<R> *Qb) {
  Mat<R, 2, 2> Cm = adjmul<R, 2, 4, 2>(Wa, Vb);
  double sid = val(abs2(Cm(0, 0)) + abs2(Cm(1, 1)));
  double ssw = val(abs2(Cm(0, 1)) + abs2(Cm(1, 0)));
  if (ssw > sid) {
    for (int i = 0; i < 4; ++i) { std::swap(Vb(i, 0), Vb(i, 1)); std::swap(Wb(i, 0), Wb(i, 1)); }
    std::swap(Qb[0], Qb[1]);
  }
}

// _modes_stack_core -- mod_ModeSolver.jl:358-392 (the objectid cache only
// avoids recomputing identical solves, so it does not change results)
template <class R> bool modes_stack(const std::vector<Layer<R>> &st, R omega, R kx, R ky, std::vector<Modes<R>> &out) {
  int N = (int)st.size();
  out.resize(N);
  Modes<R> M0;
  if (!modes_layer(st[0], omega, kx, ky, M0)) return false;
  out[0] = M0;
  Mat<R, 4, 2> W0f = M0.Wf, W0b = M0.Wb;
  for (int j = 1; j < N; ++j) {
    Modes<R> Mj;
    if (!modes_layer(st[j], omega, kx, ky, Mj)) return false;
    match_modes(W0f, Mj.Vf, Mj.Wf, Mj.Qf);
    match_modes(W0b, Mj.Vb, Mj.Wb, Mj.Qb);
    out[j] = Mj;
    W0f = Mj.Wf; W0b = Mj.Wb;
  }
  return true;
}

// ============================================================================
// S-matrix cascade (mod_ScatteringMatrix.jl)
// ============================================================================
// _S_Interface_Matrix -- :27-47
template <class R> Mat<R, 4, 4> S_interface(const Mat<R, 4, 2> &VfA, const Mat<R, 4, 2> &VbA,
                                            const Mat<R, 4, 2> &VfB, const Mat<R, 4, 2> &VbB) {
  Mat<R, 4, 4> L, Rh, S;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 2; ++j) {
      L(i, j) = -VfB(i, j); L(i, j + 2) = VbA(i, j);
      Rh(i, j) = -VfA(i, j); Rh(i, j + 2) = VbB(i, j);
    }
  if (!lu_solve4<R, 4>(L, Rh, S)) qr_solve4<R, 4>(L, Rh, S);
  return S;
}
// _SProp -- :50-62 (diagonal only)
template <class R> void S_prop(const Modes<R> &M, double d, cx<R> *dg) {
  R dd = R(d);
  dg[0] = cexp(times_i(M.Qf[0]) * Cr(dd));
  dg[1] = cexp(times_i(M.Qf[1]) * Cr(dd));
  dg[2] = cexp(-times_i(M.Qb[0]) * Cr(dd));
  dg[3] = cexp(-times_i(M.Qb[1]) * Cr(dd));
}
template <class R> Mat<R, 4, 4> diag4(const cx<R> *dg) {
  Mat<R, 4, 4> S;
  for (int i = 0; i < 4; ++i) S(i, i) = dg[i];
  return S;
}
// _Redheffer -- :65-85
template <class R> Mat<R, 4, 4> redheffer(const Mat<R, 4, 4> &S1, const Mat<R, 4, 4> &S2) {
  Mat<R, 2, 2> A11 = blk(S1, 0, 0), A12 = blk(S1, 0, 1), A21 = blk(S1, 1, 0), A22 = blk(S1, 1, 1);
  Mat<R, 2, 2> B11 = blk(S2, 0, 0), B12 = blk(S2, 0, 1), B21 = blk(S2, 1, 0), B22 = blk(S2, 1, 1);
  Mat<R, 2, 2> I2 = eye<R, 2>();
  Mat<R, 2, 2> M1 = sub(I2, mul(A12, B21));
  Mat<R, 2, 2> M2 = sub(I2, mul(B21, A12));
  Mat<R, 2, 2> S11 = mul(B11, solve2(M1, A11));
  Mat<R, 2, 2> S12 = add(B12, mul(B11, mul(solve2(M1, A12), B22)));
  Mat<R, 2, 2> S21 = add(A21, mul(A22, mul(solve2(M2, B21), A11)));
  Mat<R, 2, 2> S22 = mul(A22, solve2(M2, B22));
  return from_blocks(S11, S12, S21, S22);
}

template <class R> struct Smat {
  std::vector<Mat<R, 4, 4>> Sint, Sprop, SLeft, SRight;
};
// S_all_ingredients + S_all_Matrix -- :94-154
template <class R> void S_all(const std::vector<Layer<R>> &st, const std::vector<Modes<R>> &modes, Smat<R> &S) {
  int N = (int)st.size();
  S.Sprop.resize(N); S.Sint.resize(N > 1 ? N - 1 : 0); S.SLeft.resize(N); S.SRight.resize(N);
  for (int j = 0; j < N; ++j) {
    cx<R> dg[4];
    S_prop(modes[j], st[j].d, dg);
    S.Sprop[j] = diag4(dg);
  }
  for (int j = 1; j < N; ++j) S.Sint[j - 1] = S_interface(modes[j - 1].Vf, modes[j - 1].Vb, modes[j].Vf, modes[j].Vb);
  Mat<R, 4, 4> I4 = eye<R, 4>();
  S.SLeft[0] = I4;
  Mat<R, 4, 4> S0 = redheffer(I4, S.Sprop[0]);
  for (int j = 1; j < N; ++j) {
    Mat<R, 4, 4> St = redheffer(S0, S.Sint[j - 1]);
    S.SLeft[j] = St;
    S0 = redheffer(St, S.Sprop[j]);
  }
  S.SRight[N - 1] = I4;
  for (int k = N - 2; k >= 0; --k) {
    Mat<R, 4, 4> eff = redheffer(S.Sprop[k], S.Sint[k]);
    S.SRight[k] = redheffer(eff, S.SRight[k + 1]);
  }
}

// ============================================================================
// Source projection, cavity, walk, T^o (mod_SourceProj.jl, mod_FieldSolver.jl)
// ============================================================================
// Source_projection_Berreman_xyz -- mod_SourceProj.jl:34-60
template <class R> bool source_projection(R kx, R ky, R omega, const Layer<R> &L, const Modes<R> &M,
                                          Mat<R, 2, 3> &sif, Mat<R, 2, 3> &sib) {
  typedef cx<R> T;
  T etz1 = L.e(0, 2), etz2 = L.e(1, 2), ezz = L.e(2, 2);
  T mio = cx<R>(R(0.0), -omega);  // -1im * omega
  // Jt_cols = mio * [1 0 0; 0 1 0], Jz_row = mio * [0 0 1]
  Mat<R, 4, 3> J;
  T k1 = Cr(kx), k2 = Cr(ky), om = Cr(omega);
  T zero = C<R>(0.0);
  T den = ezz * om;
  // J_Et = kts * transpose(Jz_row) / (ezz*omega)
  T jz[3] = {rmul(R(0.0), mio), rmul(R(0.0), mio), mio};
  for (int q = 0; q < 3; ++q) { J(0, q) = k1 * jz[q] / den; J(1, q) = k2 * jz[q] / den; }
  // J_Ht = R2s * Jt_cols - (R2s * etz) * transpose(Jz_row) / ezz ; R2s = [0 1; -1 0]
  T jt[2][3] = {{mio, rmul(R(0.0), mio), rmul(R(0.0), mio)}, {rmul(R(0.0), mio), mio, rmul(R(0.0), mio)}};
#if ORACLE_MUTANT == 2
  T r1 = zero, r2 = zero;
  (void)etz1; (void)etz2;
#else
  T r1 = etz2, r2 = -etz1;  // R2s * etz
#endif
  for (int q = 0; q < 3; ++q) {
    T a = jt[1][q], b = -jt[0][q];  // R2s * Jt_cols
    J(2, q) = a - r1 * jz[q] / ezz;
    J(3, q) = b - r2 * jz[q] / ezz;
  }
  (void)zero;
  Mat<R, 4, 4> Vm, Wm;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 2; ++j) {
      Vm(i, j) = M.Vf(i, j); Vm(i, j + 2) = M.Vb(i, j);
      Wm(i, j) = M.Wf(i, j); Wm(i, j + 2) = M.Wb(i, j);
    }
  Mat<R, 4, 4> A = adjmul<R, 4, 4, 4>(Wm, Vm);
  Mat<R, 4, 3> B = adjmul<R, 4, 4, 3>(Wm, J);
  Mat<R, 4, 3> s;
  if (!lu_solve4<R, 3>(A, B, s)) return false;  // SingularException -> node invalid
  for (int q = 0; q < 3; ++q) {
    sif(0, q) = s(0, q); sif(1, q) = s(1, q);
    sib(0, q) = s(2, q); sib(1, q) = s(3, q);
  }
  return true;
}

// _flush_denormal -- mod_FieldSolver.jl:24-26
template <class R, int M, int N> inline Mat<R, M, N> flush(const Mat<R, M, N> &A) {
  Mat<R, M, N> B = A;
  for (int i = 0; i < M * N; ++i)
    if (val(cabs(B.a[i])) < 1e-250) B.a[i] = C<R>(0.0);
  return B;
}
// _cond2x2 / _cond2x2_eq -- mod_FieldSolver.jl:33-53
template <class R> double cond2x2(const Mat<R, 2, 2> &A) {
  Mat<R, 2, 2> G = adjmul<R, 2, 2, 2>(A, A);
  R tr = (G(0, 0) + G(1, 1)).re;
  R det = (G(0, 0) * G(1, 1) - G(0, 1) * G(1, 0)).re;
  R dsc = tr * tr - R(4.0) * det;
  R disc = r_sqrt(R(std::max(val(dsc), 0.0)));
  R l1 = (tr + disc) / R(2.0), l2 = (tr - disc) / R(2.0);
  if (val(l2) <= 0) return INFINITY;
  return val(r_sqrt(l1 / l2));
}
template <class R> double cond2x2_eq(const Mat<R, 2, 2> &A) {
  R n1 = r_sqrt(abs2(A(0, 0)) + abs2(A(1, 0)));
  R n2 = r_sqrt(abs2(A(0, 1)) + abs2(A(1, 1)));
  if (!(val(n1) > 0 && val(n2) > 0 && std::isfinite(val(n1)) && std::isfinite(val(n2)))) return INFINITY;
  cx<R> dg[2] = {Cr<R>(R(1.0) / n1), Cr<R>(R(1.0) / n2)};
  return cond2x2(muld(A, dg));
}
// FieldSolver._SProp(d, Q) = exp(i Q d) -- mod_FieldSolver.jl:56-61
template <class R> inline void sprop2(double d, const cx<R> *Q, cx<R> *out) {
  out[0] = cexp(times_i(Q[0]) * Cr(R(d)));
  out[1] = cexp(times_i(Q[1]) * Cr(R(d)));
}

template <class R> struct Cavity {
  Mat<R, 2, 2> RL, RR;
  Mat<R, 4, 4> RR_right;
  bool has_right = false;
};
// _R_cavity -- mod_FieldSolver.jl:72-89  (layer is 0-based here)
template <class R> Cavity<R> R_cavity(const std::vector<Layer<R>> &st, int layer, const std::vector<Modes<R>> &modes,
                                      const Smat<R> &S) {
  int N = (int)st.size();
  double d = st[layer].d;
  Cavity<R> cv;
  cv.RL = blk(S.SLeft[layer], 0, 1);
  if (layer == N - 1) { cv.RR = zeros<R, 2, 2>(); return cv; }
  cx<R> dg[4];
  sprop2(-d, modes[layer].Qf, dg);
  sprop2(d, modes[layer].Qb, dg + 2);
  cv.RR_right = redheffer(diag4(dg), S.SRight[layer]);
  cv.has_right = true;
  cv.RR = blk(cv.RR_right, 1, 0);
  return cv;
}
// _cavity_resummation -- mod_FieldSolver.jl:142-152
template <class R, int NC> void cavity_resummation(const Mat<R, 2, NC> &Sp, const Mat<R, 2, NC> &Sm, const Cavity<R> &cv,
                                                   const Modes<R> &M, double d, int N, int layer,
                                                   Mat<R, 2, NC> &Fp, Mat<R, 2, NC> &Fm) {
  if (layer == N - 1) { Fp = zeros<R, 2, NC>(); Fm = Sm; return; }
  cx<R> PF[2], PB[2];
  sprop2(d, M.Qf, PF);
  sprop2(-d, M.Qb, PB);
  Mat<R, 2, 2> I2 = eye<R, 2>();
  Mat<R, 2, 2> A = sub(I2, mul(muld(dmul(PF, cv.RL), PB), cv.RR));
  Mat<R, 2, NC> rhs = add(Sp, dmul(PF, mul(cv.RL, Sm)));
  Fp = solve2(A, rhs);
  Fm = add(Sm, dmul(PB, mul(cv.RR, Fp)));
}
// _walk_stack -- mod_FieldSolver.jl:191-290.  want[k] marks the layers whose
// amplitudes are needed ("layers" keyword); nullptr = all.
template <class R, int NC> void walk_stack(const std::vector<Layer<R>> &st, int layer, const Mat<R, 2, NC> &Fp,
                                           const Mat<R, 2, NC> &Fm, const Cavity<R> &cv,
                                           const std::vector<Modes<R>> &modes, const Smat<R> &S,
                                           const std::vector<char> *want, std::vector<Mat<R, 2, NC>> &Ef,
                                           std::vector<Mat<R, 2, NC>> &Eb, double tol = 1e-8) {
  int N = (int)st.size();
  Mat<R, 2, NC> Z = zeros<R, 2, NC>();
  Ef.assign(N, Z); Eb.assign(N, Z);
  Mat<R, 2, NC> Ef_left = mul(cv.RL, Fm);
  bool want_self = !want || (*want)[layer];
  bool need_right = !want, need_left = !want;
  if (want) {
    for (int k = 0; k < N; ++k) {
      if ((*want)[k] && k > layer) need_right = true;
      if ((*want)[k] && k < layer) need_left = true;
    }
    if (want_self && layer < N - 1) need_right = true;
  }
  if (layer < N - 1 && need_right) {
    Mat<R, 2, 2> S11 = blk(cv.RR_right, 0, 0), S12 = blk(cv.RR_right, 0, 1);
    Mat<R, 2, NC> cm = Z;
    Mat<R, 2, NC> cp = add(mul(S11, Fp), mul(S12, cm));
    Ef[N - 1] = add(Ef[N - 1], flush(cp));
    Eb[N - 1] = add(Eb[N - 1], flush(cm));
    for (int k = layer + 1; k <= N - 2; ++k) {
      const Mat<R, 4, 4> &Sk = S.SRight[k];
      Mat<R, 2, 2> S11k = blk(Sk, 0, 0), S12k = blk(Sk, 0, 1), S21k = blk(Sk, 1, 0), S22k = blk(Sk, 1, 1);
      Mat<R, 2, NC> bp;
      if (cond2x2_eq(S11k) > 1.0 / tol) bp = Z;
      else bp = solve2(S11k, sub(cp, mul(S12k, cm)));
      Mat<R, 2, NC> bm = add(mul(S21k, bp), mul(S22k, cm));
      Ef[k] = add(Ef[k], flush(bp));
      Eb[k] = add(Eb[k], flush(bm));
    }
  }
  if (layer > 0 && need_left) {
    const Mat<R, 4, 4> &Sj = S.SLeft[layer];
    Mat<R, 2, 2> S22 = blk(Sj, 1, 1);
    Mat<R, 2, NC> ap = Z;
    Mat<R, 2, NC> am = mul(S22, Fm);
    Ef[0] = add(Ef[0], flush(ap));
    Eb[0] = add(Eb[0], flush(am));
    for (int k = layer - 1; k >= 1; --k) {
      const Mat<R, 4, 4> &Sk = S.SLeft[k];
      Mat<R, 2, 2> S11k = blk(Sk, 0, 0), S12k = blk(Sk, 0, 1), S21k = blk(Sk, 1, 0), S22k = blk(Sk, 1, 1);
      Mat<R, 2, NC> bm;
      if (cond2x2_eq(S22k) > 1.0 / tol) bm = Z;
      else bm = solve2(S22k, sub(am, mul(S21k, ap)));
      Mat<R, 2, NC> bp = add(mul(S11k, ap), mul(S12k, bm));
      Ef[k] = add(Ef[k], flush(bp));
      Eb[k] = add(Eb[k], flush(bm));
    }
  }
  for (int k = 0; k <= N - 2; ++k) {
    Mat<R, 2, NC> Ef_far;
    if (k == layer) Ef_far = Fp;
    else {
      cx<R> ph[2];
      sprop2(st[k].d, modes[k].Qf, ph);
      Ef_far = dmul(ph, Ef[k]);
    }
    Mat<R, 2, NC> Eb_next = (k == layer - 1) ? Fm : Eb[k + 1];
    Mat<R, 2, 2> S21k = blk(S.Sint[k], 1, 0), S22k = blk(S.Sint[k], 1, 1);
    Eb[k] = flush(add(mul(S21k, Ef_far), mul(S22k, Eb_next)));
  }
  Ef[layer] = flush(Ef_left);
  if (layer == N - 1) Eb[N - 1] = flush(Fm);
}
// _cavity_transfer_columns -- mod_FieldSolver.jl:320-341.  obs are 0-based.
template <class R> void cavity_transfer_columns(const std::vector<Layer<R>> &st, int layer,
                                                const std::vector<Modes<R>> &modes, const Smat<R> &S,
                                                const std::vector<int> &obs, std::vector<Mat<R, 4, 4>> &T) {
  int N = (int)st.size();
  Cavity<R> cv = R_cavity(st, layer, modes, S);
  Mat<R, 2, 4> Ep = zeros<R, 2, 4>(), Em = zeros<R, 2, 4>();
  if (layer < N - 1) { Ep(0, 0) = C<R>(1.0); Ep(1, 1) = C<R>(1.0); }
  Em(0, 2) = C<R>(1.0); Em(1, 3) = C<R>(1.0);
  Mat<R, 2, 4> Fp, Fm;
  cavity_resummation<R, 4>(Ep, Em, cv, modes[layer], st[layer].d, N, layer, Fp, Fm);
  std::vector<char> want(N, 0);
  bool any = false;
  for (int o : obs)
    if (o != layer) { want[o] = 1; any = true; }
  std::vector<Mat<R, 2, 4>> Ef, Eb;
  if (any) walk_stack<R, 4>(st, layer, Fp, Fm, cv, modes, S, &want, Ef, Eb);
  T.resize(obs.size());
  for (size_t i = 0; i < obs.size(); ++i) {
    int o = obs[i];
    Mat<R, 2, 4> a, b;
    if (o == layer) { a = mul(cv.RL, Fm); b = mul(cv.RR, Fp); }
    else { a = Ef[o]; b = Eb[o]; }
    for (int j = 0; j < 4; ++j) {
      T[i](0, j) = a(0, j); T[i](1, j) = a(1, j);
      T[i](2, j) = b(0, j); T[i](3, j) = b(1, j);
    }
  }
}

// _mode_field_core -- mod_CavityTransfer.jl:87-105 ; fills one obs basis
// (Exf[2], Eyf[2], Ezf[2], Exb[2], Eyb[2], Ezb[2], KZf[2], KZb[2]) as in
// _fill_obs_basis! (:118-130): a mode whose 8 values are not all finite is left 0.
template <class R> void fill_obs_basis(const Layer<R> &L, const Modes<R> &M, R omega, R kx, R ky, cx<R> *B16) {
  cx<R> ezt1 = L.e(2, 0), ezt2 = L.e(2, 1), ezz = L.e(2, 2);
  cx<R> iw = C<R>(1.0) / Cr(omega);
  cx<R> k1 = Cr(kx), k2 = Cr(ky);
  for (int m = 0; m < 2; ++m) {
    cx<R> vals[8];
    for (int dir = 0; dir < 2; ++dir) {
      const Mat<R, 4, 2> &V = dir == 0 ? M.Vf : M.Vb;
      cx<R> kz = dir == 0 ? M.Qf[m] : M.Qb[m];
      cx<R> ex = V(0, m), ey = V(1, m);
      cx<R> ez = -((ezt1 * V(0, m) + ezt2 * V(1, m) + iw * (k1 * V(3, m) - k2 * V(2, m))) / ezz);
      vals[dir * 3 + 0] = ex; vals[dir * 3 + 1] = ey; vals[dir * 3 + 2] = ez;
      vals[6 + dir] = kz;
    }
    bool ok = true;
    for (int i = 0; i < 8; ++i) ok = ok && isfin(vals[i]);
    if (!ok) continue;
    B16[0 + m] = vals[0]; B16[2 + m] = vals[1]; B16[4 + m] = vals[2];
    B16[6 + m] = vals[3]; B16[8 + m] = vals[4]; B16[10 + m] = vals[5];
    B16[12 + m] = vals[6]; B16[14 + m] = vals[7];
  }
}

// One Green k-node -- mod_CavityTransfer.jl:224-251 (1D) / :339-359, 383-388 (2D)
template <class R> struct NodeOut {
  bool valid = false;
  std::vector<char> Tvalid;
  std::vector<Mat<R, 4, 4>> T;
  cx<R> Qf[2], Qb[2];
  Mat<R, 2, 3> sf, sb;
  std::vector<cx<R>> basis;  // nobs * 16
};
template <class R> void node_solve(const std::vector<Layer<R>> &st, R omega, R kx, R ky, int src,
                                   const std::vector<int> &obs, NodeOut<R> &out) {
  size_t no = obs.size();
  out.valid = false;
  out.Tvalid.assign(no, 0);
  out.T.assign(no, zeros<R, 4, 4>());
  out.basis.assign(no * 16, C<R>(0.0));
  out.Qf[0] = out.Qf[1] = out.Qb[0] = out.Qb[1] = C<R>(0.0);
  out.sf = zeros<R, 2, 3>(); out.sb = zeros<R, 2, 3>();
  std::vector<Modes<R>> modes;
  if (!modes_stack(st, omega, kx, ky, modes)) return;
  Smat<R> S;
  S_all(st, modes, S);
  std::vector<Mat<R, 4, 4>> Tk;
  cavity_transfer_columns(st, src, modes, S, obs, Tk);
  Mat<R, 2, 3> sif, sib;
  if (!source_projection(kx, ky, omega, st[src], modes[src], sif, sib)) return;
  if (!(allfinite(sif) && allfinite(sib))) return;
  for (size_t i = 0; i < no; ++i)
    if (allfinite(Tk[i])) { out.T[i] = Tk[i]; out.Tvalid[i] = 1; }
  out.Qf[0] = modes[src].Qf[0]; out.Qf[1] = modes[src].Qf[1];
  out.Qb[0] = modes[src].Qb[0]; out.Qb[1] = modes[src].Qb[1];
  out.sf = sif;
  out.sb = neg(sib);
  for (size_t i = 0; i < no; ++i) fill_obs_basis(st[obs[i]], modes[obs[i]], omega, kx, ky, &out.basis[i * 16]);
  out.valid = true;
}

// ============================================================================
// Synthesis helpers (mod_ClosedFormField.jl)
// ============================================================================
// _phase -- :59-69
template <class R> inline cx<R> phase(cx<R> kz, double z) {
  R a = -(kz.im * R(z));
  R b = kz.re * R(z);
  double av = val(a);
  if (!std::isfinite(av) || !std::isfinite(val(b)) || av > 700 || av < -700) return C<R>(0.0);
  return cexp(cx<R>(a, b));
}
template <class R> inline cx<R> finite_or_zero(cx<R> z) { return isfin(z) ? z : C<R>(0.0); }

// _boxphase (mod_ClosedFormField.jl:86-91): the box-averaged exponential in
// face-difference form, [e^{iQ(l+dz/2)} - e^{iQ(l-dz/2)}] / (iQ dz), with the
// 4th-order Taylor guard for |Q dz| < 1e-3; dz = 0 is the clamped phase.
template <class R> inline cx<R> boxphase(cx<R> Q, double l, double dz) {
  if (dz == 0.0) return phase(Q, l);
  const cx<R> y = rmul(R(dz), Q);
  if (val(cabs(y)) < 1e-3) {
    const cx<R> y2 = y * y;
    return phase(Q, l) * (C<R>(1.0) - rmul(R(1.0 / 24.0), y2) + rmul(R(1.0 / 1920.0), y2 * y2));
  }
  return (phase(Q, l + dz / 2) - phase(Q, l - dz / 2)) / times_i(y);
}
// Julia's expm1(::Complex): expm1(x) - 2 e^x sin^2(t/2) + i e^x sin t
template <class R> inline cx<R> cexpm1(cx<R> z) {
  const double x = val(z.re), t = val(z.im);
  const double em1 = std::expm1(x);
  if (t == 0.0) return C<R>(em1, t);
  const double er = em1 + 1.0, sh = std::sin(0.5 * t);
  return C<R>(em1 - 2.0 * er * sh * sh, er * std::sin(t));
}
// _ramp (:790-796): expm1(iQa) / (iQ dz), regular as Q -> 0
template <class R> inline cx<R> ramp(cx<R> Q, double a, double dz) {
  if (val(cabs(Q)) * a < 1e-12) return C<R>(a / dz);
  const cx<R> y = times_i(rmul(R(a), Q));
  const cx<R> den = times_i(rmul(R(dz), Q));
  if (val(y.re) < -700) return C<R>(-1.0) / den;
  return cexpm1(y) / den;
}

// Spectral dyadic of one node for one (zsrc, zobs) pair: the strength of
// _strengths_row! (:616-659) divided by its node weight,
//   g[c,q] = sum_m E_c,f fz(Af_r Phi_f + Afd) + E_c,b fz(Ab_r Phi_b + Abd),
// with amplitude_at (:539-566), direct_amplitude (:577-602, point source),
// _obs_phases (:326-335); zpos convention :72-73.  Output g[c + 3q].
// terms (_term_amplitudes, :143-160): 0 :all, 1 :direct, 2 :scattered,
// 3 :toeplitz (direct + f->f, b->b reflected), 4 :hankel (b->f, f->b).
template <class R> void node_ghat(const NodeOut<R> &nd, int obs_idx, bool same_layer, int src_layer0, double d_src,
                                  double d_obs, double zsrc_local, double zobs, cx<R> *g, int terms = 0,
                                  double dz = 0.0, bool ez_local = true, cx<R> ezz_src = cx<R>()) {
  for (int i = 0; i < 9; ++i) g[i] = C<R>(0.0);
  double zpos = src_layer0 == 0 ? d_src - zsrc_local : zsrc_local;
  cx<R> Af[2][3], Ab[2][3], Afd[2][3], Abd[2][3];
  for (int m = 0; m < 2; ++m)
    for (int q = 0; q < 3; ++q) Af[m][q] = Ab[m][q] = Afd[m][q] = Abd[m][q] = C<R>(0.0);
  if (nd.valid && nd.Tvalid[obs_idx]) {
    cx<R> Pf[2], Pb[2];
    for (int m = 0; m < 2; ++m) {
      if (dz > 0.0) {  // _seed_phases of a box (:106-111)
        Pf[m] = boxphase(nd.Qf[m], d_src - zpos, dz);
        Pb[m] = boxphase(nd.Qb[m], -zpos, dz);
      } else {
        Pf[m] = cexp(times_i(nd.Qf[m]) * Cr(R(d_src - zpos)));
        Pb[m] = cexp(-times_i(nd.Qb[m]) * Cr(R(zpos)));
      }
    }
    const Mat<R, 4, 4> &T = nd.T[obs_idx];
    // seed columns feeding the forward (rows 0-1) and backward (rows 2-3)
    // amplitudes: seeds=:f keeps columns 0-1, seeds=:b columns 2-3
    bool useF[4] = {true, true, true, true}, useB[4] = {true, true, true, true};
    if (terms == 1)
      for (int k = 0; k < 4; ++k) useF[k] = useB[k] = false;
    if (terms == 3)
      for (int k = 0; k < 4; ++k) { useF[k] = k < 2; useB[k] = k >= 2; }
    if (terms == 4)
      for (int k = 0; k < 4; ++k) { useF[k] = k >= 2; useB[k] = k < 2; }
    for (int q = 0; q < 3; ++q) {
      cx<R> u[4] = {nd.sf(0, q) * Pf[0], nd.sf(1, q) * Pf[1], nd.sb(0, q) * Pb[0], nd.sb(1, q) * Pb[1]};
      cx<R> v[4];
      for (int i = 0; i < 4; ++i) {
        const bool *use = i < 2 ? useF : useB;
        cx<R> s = C<R>(0.0);
        bool first = true;
        for (int k = 0; k < 4; ++k)
          if (use[k]) {
            s = first ? T(i, k) * u[k] : s + T(i, k) * u[k];
            first = false;
          }
        v[i] = s;
      }
      Af[0][q] = v[0]; Af[1][q] = v[1]; Ab[0][q] = v[2]; Ab[1][q] = v[3];
    }
  }
  bool inside = false;
  if (same_layer && nd.valid && (terms == 0 || terms == 1 || terms == 3)) {
    double zeta = zobs - zpos;
    if (dz > 0.0 && std::fabs(zeta) < dz / 2) {
      // interior_slab_at (:815-843): face ramps, already at z_obs
      inside = true;
      const double a = std::min(std::max(zobs - (zpos - dz / 2), 0.0), dz);
      const double b = std::min(std::max((zpos + dz / 2) - zobs, 0.0), dz);
      for (int m = 0; m < 2; ++m) {
        const cx<R> rf = ramp(nd.Qf[m], a, dz), rb = ramp(-nd.Qb[m], b, dz);
        for (int q = 0; q < 3; ++q) {
          Afd[m][q] = nd.sf(m, q) * rf;
          Abd[m][q] = nd.sb(m, q) * rb;
        }
      }
    } else {
      // _direct_branch (:208-214): point -> strict sign of zeta, box -> |zeta| >= dz/2
      bool fwd = dz > 0.0 ? zeta >= dz / 2 : zeta > 0, bwd = dz > 0.0 ? zeta <= -dz / 2 : zeta < 0;
      for (int m = 0; m < 2; ++m) {
        if (fwd) {
          cx<R> e = boxphase(nd.Qf[m], zeta, dz);
          for (int q = 0; q < 3; ++q) Afd[m][q] = nd.sf(m, q) * e;
        } else if (bwd) {
          cx<R> e = boxphase(nd.Qb[m], zeta, dz);
          for (int q = 0; q < 3; ++q) Abd[m][q] = nd.sb(m, q) * e;
        }
      }
    }
  }
  const cx<R> *B = &nd.basis[obs_idx * 16];
  for (int m = 0; m < 2; ++m) {
    cx<R> pf = phase(B[12 + m], zobs), pb = phase(B[14 + m], zobs - d_obs);
    for (int q = 0; q < 3; ++q) {
      cx<R> af = finite_or_zero(Af[m][q] * pf + Afd[m][q]);
      cx<R> ab = finite_or_zero(Ab[m][q] * pb + Abd[m][q]);
      g[0 + 3 * q] = g[0 + 3 * q] + (B[0 + m] * af + B[6 + m] * ab);
      g[1 + 3 * q] = g[1 + 3 * q] + (B[2 + m] * af + B[8 + m] * ab);
      g[2 + 3 * q] = g[2 + 3 * q] + (B[4 + m] * af + B[10 + m] * ab);
    }
  }
  // the local Ez of a box inside its own slab (ez_local_density, :861-862):
  // -p_z / (eps_zz dz), on the zz entry of valid nodes
  if (inside && ez_local) g[8] = g[8] - C<R>(1.0) / (rmul(R(dz), ezz_src));
}

// Modal_amp -- mod_FieldSolver.jl:356-363, with one ModalSeed (z, sif, sib)
// and _seed_from_sources (:106-127).  Used by tests only.
template <class R> void modal_amp(const std::vector<Layer<R>> &st, const std::vector<Modes<R>> &modes, const Smat<R> &S,
                                  int layer, double z, const cx<R> *sif, const cx<R> *sib, std::vector<Mat<R, 2, 1>> &Ef,
                                  std::vector<Mat<R, 2, 1>> &Eb, Mat<R, 2, 1> *FpOut = nullptr,
                                  Mat<R, 2, 1> *FmOut = nullptr, Cavity<R> *cvOut = nullptr) {
  int N = (int)st.size();
  Cavity<R> cv = R_cavity(st, layer, modes, S);
  double d = st[layer].d;
  double zpos = layer == 0 ? d - z : z;
  Mat<R, 2, 1> Sp = zeros<R, 2, 1>(), Sm = zeros<R, 2, 1>();
  if (layer < N - 1) {
    cx<R> ph[2];
    sprop2(d - zpos, modes[layer].Qf, ph);
    Sp(0, 0) = sif[0] * ph[0]; Sp(1, 0) = sif[1] * ph[1];
  }
  cx<R> pb[2];
  sprop2(-zpos, modes[layer].Qb, pb);
  Sm(0, 0) = sib[0] * pb[0]; Sm(1, 0) = sib[1] * pb[1];
  Mat<R, 2, 1> Fp, Fm;
  cavity_resummation<R, 1>(Sp, Sm, cv, modes[layer], d, N, layer, Fp, Fm);
  walk_stack<R, 1>(st, layer, Fp, Fm, cv, modes, S, nullptr, Ef, Eb);
  if (FpOut) *FpOut = Fp;
  if (FmOut) *FmOut = Fm;
  if (cvOut) *cvOut = cv;
}

template <class R> std::vector<Layer<R>> make_stack(int n, const double *eps, const double *mu, const double *d,
                                                    const int *ek, const int *mk) {
  std::vector<Layer<R>> st(n);
  for (int l = 0; l < n; ++l) {
    for (int i = 0; i < 9; ++i) {
      st[l].eps[i] = C<R>(eps[18 * l + 2 * i], eps[18 * l + 2 * i + 1]);
      st[l].mu[i] = C<R>(mu[18 * l + 2 * i], mu[18 * l + 2 * i + 1]);
    }
    st[l].d = d[l];
    st[l].ek = ek[l];
    st[l].mk = mk[l];
  }
  return st;
}

template <class R> inline void put(double *o, cx<R> z) { o[0] = val(z.re); o[1] = val(z.im); }
template <class R> inline cx<R> get(const double *o) { return C<R>(o[0], o[1]); }

// Per-node record in the oracle's flat cache, complex entries:
//   [0:4)   Qf1 Qf2 Qb1 Qb2
//   [4:10)  sf (2x3 col-major)      [10:16) sb
//   per obs i: base = 16 + 32 i:  T (16, col-major 4x4), basis (16)
// plus flags valid (int) and Tvalid[nobs] (int).
inline int rec_len(int nobs) { return 16 + 32 * nobs; }

template <class R> void write_rec(const NodeOut<R> &nd, int nobs, double *rec, int *valid, int *Tvalid) {
  put(rec + 0, nd.Qf[0]); put(rec + 2, nd.Qf[1]); put(rec + 4, nd.Qb[0]); put(rec + 6, nd.Qb[1]);
  for (int i = 0; i < 6; ++i) { put(rec + 8 + 2 * i, nd.sf.a[i]); put(rec + 20 + 2 * i, nd.sb.a[i]); }
  for (int o = 0; o < nobs; ++o) {
    double *b = rec + 2 * (16 + 32 * o);
    for (int i = 0; i < 16; ++i) put(b + 2 * i, nd.T[o].a[i]);
    for (int i = 0; i < 16; ++i) put(b + 32 + 2 * i, nd.basis[16 * o + i]);
    Tvalid[o] = nd.Tvalid[o];
  }
  *valid = nd.valid;
}
template <class R> void read_rec(NodeOut<R> &nd, int nobs, const double *rec, int valid, const int *Tvalid) {
  nd.valid = valid != 0;
  nd.Tvalid.assign(nobs, 0);
  nd.T.assign(nobs, zeros<R, 4, 4>());
  nd.basis.assign(16 * nobs, C<R>(0.0));
  nd.Qf[0] = get<R>(rec + 0); nd.Qf[1] = get<R>(rec + 2); nd.Qb[0] = get<R>(rec + 4); nd.Qb[1] = get<R>(rec + 6);
  for (int i = 0; i < 6; ++i) { nd.sf.a[i] = get<R>(rec + 8 + 2 * i); nd.sb.a[i] = get<R>(rec + 20 + 2 * i); }
  for (int o = 0; o < nobs; ++o) {
    const double *b = rec + 2 * (16 + 32 * o);
    for (int i = 0; i < 16; ++i) nd.T[o].a[i] = get<R>(b + 2 * i);
    for (int i = 0; i < 16; ++i) nd.basis[16 * o + i] = get<R>(b + 32 + 2 * i);
    nd.Tvalid[o] = (char)Tvalid[o];
  }
}

// _σmax -- mod_KSpaceFeatures.jl:91-95: the largest singular value of a 2x2
// from G = A'A (trace t, determinant dt).
template <class R> double sigma_max2(const Mat<R, 2, 2> &A) {
  Mat<R, 2, 2> G = adjmul<R, 2, 2, 2>(A, A);
  const double t = val((G(0, 0) + G(1, 1)).re);
  const double dt = val((G(0, 0) * G(1, 1) - G(0, 1) * G(1, 0)).re);
  return std::sqrt(std::max((t + std::sqrt(std::max(t * t - 4.0 * dt, 0.0))) / 2.0, 0.0));
}

// _cavity_response -- mod_KSpaceFeatures.jl:109-127 (source layer s 1-based):
// RL = SLeft[s][1:2, 3:4], RR = (Sint[s] * SRight[s+1])[3:4, 1:2] ("*" the
// Redheffer product), r = max(sigma(RL), sigma(RR)) and, for s > 1, also
// sigma(inv(I - PhiF RL PhiB RR)) with PhiF = e^{i Qf d}, PhiB = e^{-i Qb d};
// 0 where the mode solve fails or r is not finite.
template <class R> double cavity_response(const std::vector<Layer<R>> &st, R omega, int s, R kx, R ky) {
  std::vector<Modes<R>> M;
  if (!modes_stack(st, omega, kx, ky, M)) return 0.0;
  Smat<R> S;
  S_all(st, M, S);
  const Mat<R, 4, 4> X = redheffer(S.Sint[s - 1], S.SRight[s]);
  const Mat<R, 2, 2> RR = blk(X, 1, 0), RL = blk(S.SLeft[s - 1], 0, 1);
  double r = std::max(sigma_max2(RL), sigma_max2(RR));
  if (s > 1) {
    const double d = st[s - 1].d;
    cx<R> pf[2], pb[2];
    sprop2(d, M[s - 1].Qf, pf);
    sprop2(-d, M[s - 1].Qb, pb);
    const Mat<R, 2, 2> A = mul<R, 2, 2, 2>(dmul(pf, RL), dmul(pb, RR));
    Mat<R, 2, 2> Mm = sub(eye<R, 2>(), A);
    const cx<R> det = Mm(0, 0) * Mm(1, 1) - Mm(0, 1) * Mm(1, 0);
    Mat<R, 2, 2> Mi;
    Mi(0, 0) = Mm(1, 1) / det;
    Mi(1, 0) = -Mm(1, 0) / det;
    Mi(0, 1) = -Mm(0, 1) / det;
    Mi(1, 1) = Mm(0, 0) / det;
    r = std::max(r, sigma_max2(Mi));
  }
  return std::isfinite(r) ? r : 0.0;
}

}  // namespace ora


// ============================================================================
// C API (ctypes) -- test infrastructure
// Stack arrays: eps, mu: n*18 doubles (9 interleaved complex, column-major
// 3x3 per layer); d: n; ek, mk: n.  Layer indices are 1-based (Julia's).
// ============================================================================
using namespace ora;

extern "C" {

int ora_eigen4(const double *A, double *Q, double *V) {
  Mat<double, 4, 4> M, Vm;
  for (int i = 0; i < 16; ++i) M.a[i] = C<double>(A[2 * i], A[2 * i + 1]);
  cx<double> q[4];
  bool ok = eigen4(M, q, Vm);
  for (int i = 0; i < 4; ++i) put(Q + 2 * i, q[i]);
  for (int i = 0; i < 16; ++i) put(V + 2 * i, Vm.a[i]);
  return ok ? 0 : 1;
}

// [L] \ [Rh] as _S_Interface_Matrix solves it (mod_ScatteringMatrix.jl:35-44):
// LU with partial pivoting, the pivoted-QR minimum-norm solve on an exactly
// zero pivot.  Returns 1 when the QR fallback was taken, 0 otherwise.
int ora_interface_solve(const double *L, const double *Rh, double *S) {
  Mat<double, 4, 4> A, B, X;
  for (int i = 0; i < 16; ++i) {
    A.a[i] = C<double>(L[2 * i], L[2 * i + 1]);
    B.a[i] = C<double>(Rh[2 * i], Rh[2 * i + 1]);
  }
  int fb = 0;
  if (!lu_solve4<double, 4>(A, B, X)) {
    qr_solve4<double, 4>(A, B, X);
    fb = 1;
  }
  for (int i = 0; i < 16; ++i) put(S + 2 * i, X.a[i]);
  return fb;
}

// The pivoted-QR minimum-norm solve alone (qr_solve4).
int ora_qr_solve4(const double *L, const double *Rh, double *S) {
  Mat<double, 4, 4> A, B, X;
  for (int i = 0; i < 16; ++i) {
    A.a[i] = C<double>(L[2 * i], L[2 * i + 1]);
    B.a[i] = C<double>(Rh[2 * i], Rh[2 * i + 1]);
  }
  qr_solve4<double, 4>(A, B, X);
  for (int i = 0; i < 16; ++i) put(S + 2 * i, X.a[i]);
  return 0;
}

int ora_berreman(const double *eps, const double *mu, double omega, double kx, double ky, double *B) {
  int ek = EPS_GENERAL, mk = MU_GENERAL;
  double d = 0;
  auto st = make_stack<double>(1, eps, mu, &d, &ek, &mk);
  Mat<double, 4, 4> M = berreman_core(st[0], omega, kx, ky);
  for (int i = 0; i < 16; ++i) put(B + 2 * i, M.a[i]);
  return 0;
}

// Raw (Q, V) of _eigen_QV and the sorted modes of _modes_layer for one layer.
int ora_modes_layer(const double *eps, const double *mu, int ek, int mk, double omega, double kx, double ky,
                    double *Qraw, double *Vraw, double *modes36) {
  double d = 0;
  auto st = make_stack<double>(1, eps, mu, &d, &ek, &mk);
  cx<double> q[4];
  Mat<double, 4, 4> V;
  if (!eigen_QV(st[0], omega, kx, ky, q, V)) return 1;
  for (int i = 0; i < 4; ++i) put(Qraw + 2 * i, q[i]);
  for (int i = 0; i < 16; ++i) put(Vraw + 2 * i, V.a[i]);
  Modes<double> M;
  if (!modes_layer(st[0], omega, kx, ky, M)) return 2;
  for (int i = 0; i < 8; ++i) {
    put(modes36 + 2 * i, M.Vf.a[i]); put(modes36 + 16 + 2 * i, M.Vb.a[i]);
    put(modes36 + 32 + 2 * i, M.Wf.a[i]); put(modes36 + 48 + 2 * i, M.Wb.a[i]);
  }
  put(modes36 + 64, M.Qf[0]); put(modes36 + 66, M.Qf[1]); put(modes36 + 68, M.Qb[0]); put(modes36 + 70, M.Qb[1]);
  return 0;
}

// Modes_Stack + S_all_Matrix for one node.  modes: n*36 complex
// (Vf Vb Wf Wb as 4x2 col-major, then Qf, Qb); Sint (n-1)*16; Sprop, SLeft,
// SRight n*16 complex each.
int ora_stack_smatrix(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                      double omega, double kx, double ky, double *modes, double *Sint, double *Sprop, double *SLeft,
                      double *SRight) {
  auto st = make_stack<double>(n, eps, mu, d, ek, mk);
  std::vector<Modes<double>> M;
  if (!modes_stack(st, omega, kx, ky, M)) return 1;
  Smat<double> S;
  S_all(st, M, S);
  for (int l = 0; l < n; ++l) {
    double *o = modes + 72 * l;
    for (int i = 0; i < 8; ++i) {
      put(o + 2 * i, M[l].Vf.a[i]); put(o + 16 + 2 * i, M[l].Vb.a[i]);
      put(o + 32 + 2 * i, M[l].Wf.a[i]); put(o + 48 + 2 * i, M[l].Wb.a[i]);
    }
    put(o + 64, M[l].Qf[0]); put(o + 66, M[l].Qf[1]); put(o + 68, M[l].Qb[0]); put(o + 70, M[l].Qb[1]);
    for (int i = 0; i < 16; ++i) {
      put(Sprop + 32 * l + 2 * i, S.Sprop[l].a[i]);
      put(SLeft + 32 * l + 2 * i, S.SLeft[l].a[i]);
      put(SRight + 32 * l + 2 * i, S.SRight[l].a[i]);
      if (l < n - 1) put(Sint + 32 * l + 2 * i, S.Sint[l].a[i]);
    }
  }
  return 0;
}

// Batched node solve (the Green k-node loop, mod_CavityTransfer.jl:224/380).
// rec: nn * rec_len(nobs) complex; valid: nn; Tvalid: nn*nobs.
int ora_nodes_batch(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                    double omega, const double *kx, const double *ky, int64_t nn, int src, const int *obs, int nobs,
                    double *rec, int *valid, int *Tvalid, int nthreads) {
  auto st = make_stack<double>(n, eps, mu, d, ek, mk);
  std::vector<int> ob(nobs);
  for (int i = 0; i < nobs; ++i) ob[i] = obs[i] - 1;
  int rl = rec_len(nobs);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t k = 0; k < nn; ++k) {
    NodeOut<double> nd;
    node_solve(st, omega, kx[k], ky[k], src - 1, ob, nd);
    write_rec(nd, nobs, rec + 2 * rl * k, valid + k, Tvalid + nobs * k);
  }
  return 0;
}

// Spectral dyadic g~(k; zsrc, zobs) for nodes given as an oracle cache.
// g: nn * npairs * 9 complex, layout [node][pair][c + 3q].
int ora_ghat_from_cache(int n, const double *d, int src, const int *obs, int nobs, int obs_layer, const double *rec,
                        const int *valid, const int *Tvalid, int64_t nn, int npairs, const double *zsrc,
                        const double *zobs, double *g, int nthreads) {
  int rl = rec_len(nobs);
  int oi = -1;
  for (int i = 0; i < nobs; ++i)
    if (obs[i] == obs_layer) oi = i;
  if (oi < 0) return 1;
  (void)n;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (int64_t k = 0; k < nn; ++k) {
    NodeOut<double> nd;
    read_rec(nd, nobs, rec + 2 * rl * k, valid[k], Tvalid + nobs * k);
    for (int p = 0; p < npairs; ++p) {
      cx<double> gg[9];
      node_ghat(nd, oi, obs_layer == src, src - 1, d[src - 1], d[obs_layer - 1], zsrc[p], zobs[p], gg);
      for (int i = 0; i < 9; ++i) put(g + 2 * (9 * (npairs * k + p) + i), gg[i]);
    }
  }
  return 0;
}

// Fused node solve + spectral dyadic, as timed for the CPU baseline.
// Box source of thickness dz centred at each zsrc (dz = 0: point source);
// ez_local adds the local Ez inside the box's own slab.  The box must fit in
// its layer (_check_box_in_layer, mod_ClosedFormField.jl:93-98): returns 1
// otherwise.
int ora_ghat_batch_box(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                       double omega, const double *kx, const double *ky, int64_t nn, int src, int obs_layer,
                       int npairs, const double *zsrc, const double *zobs, double *g, int nthreads, int terms,
                       double dz, int ez_local) {
  auto st = make_stack<double>(n, eps, mu, d, ek, mk);
  const double ds = d[src - 1], tol = std::sqrt(2.220446049250313e-16) * std::max(ds, 1.0);
  if (dz < 0.0) return 1;
  for (int p = 0; p < npairs && dz > 0.0; ++p) {
    const double zpos = src == 1 ? ds - zsrc[p] : zsrc[p];
    if (!(zpos - dz / 2 >= -tol && zpos + dz / 2 <= ds + tol)) return 1;
  }
  const cx<double> ezz(eps[18 * (src - 1) + 16], eps[18 * (src - 1) + 17]);
  std::vector<int> ob(1, obs_layer - 1);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t k = 0; k < nn; ++k) {
    NodeOut<double> nd;
    node_solve(st, omega, kx[k], ky[k], src - 1, ob, nd);
    for (int p = 0; p < npairs; ++p) {
      cx<double> gg[9];
      node_ghat(nd, 0, obs_layer == src, src - 1, d[src - 1], d[obs_layer - 1], zsrc[p], zobs[p], gg, terms, dz,
                ez_local != 0, ezz);
      for (int i = 0; i < 9; ++i) put(g + 2 * (9 * (npairs * k + p) + i), gg[i]);
    }
  }
  return 0;
}

int ora_ghat_batch(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                   double omega, const double *kx, const double *ky, int64_t nn, int src, int obs_layer, int npairs,
                   const double *zsrc, const double *zobs, double *g, int nthreads, int terms) {
  auto st = make_stack<double>(n, eps, mu, d, ek, mk);
  std::vector<int> ob(1, obs_layer - 1);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t k = 0; k < nn; ++k) {
    NodeOut<double> nd;
    node_solve(st, omega, kx[k], ky[k], src - 1, ob, nd);
    for (int p = 0; p < npairs; ++p) {
      cx<double> gg[9];
      node_ghat(nd, 0, obs_layer == src, src - 1, d[src - 1], d[obs_layer - 1], zsrc[p], zobs[p], gg, terms);
      for (int i = 0; i < 9; ++i) put(g + 2 * (9 * (npairs * k + p) + i), gg[i]);
    }
  }
  return 0;
}

// Batched stack S-matrix: per node, Modes_Stack (mod_ModeSolver.jl:404-413)
// and the total S of S_all_Matrix (mod_ScatteringMatrix.jl:112-154), i.e.
// SLeft[N] (= SRight[1]), with the first and last layers' modes.  Outputs per
// node (complex, may be NULL): S 16 (4x4 column-major), V1 and VN 16 each
// ([Vf Vb] 4x4 column-major), Q1 and QN 4 each (Qf1 Qf2 Qb1 Qb2); ok = 0 where
// the mode solve failed (outputs zero).
int ora_stack_S_batch(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                      double omega, const double *kx, const double *ky, int64_t nn, double *S, double *V1, double *Q1,
                      double *VN, double *QN, int *ok, int nthreads) {
  auto st = make_stack<double>(n, eps, mu, d, ek, mk);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t k = 0; k < nn; ++k) {
    std::vector<Modes<double>> M;
    bool good = modes_stack(st, omega, kx[k], ky[k], M);
    Smat<double> Sm;
    if (good) S_all(st, M, Sm);
    ok[k] = good ? 1 : 0;
    for (int i = 0; i < 16; ++i) {
      cx<double> z = good ? Sm.SLeft[n - 1].a[i] : C<double>(0.0);
      if (S) put(S + 2 * (16 * k + i), z);
    }
    const int ends[2] = {0, n - 1};
    double *Vo[2] = {V1, VN}, *Qo[2] = {Q1, QN};
    for (int e = 0; e < 2; ++e) {
      const Modes<double> *m = good ? &M[ends[e]] : nullptr;
      if (Vo[e])
        for (int i = 0; i < 8; ++i) {
          put(Vo[e] + 2 * (16 * k + i), m ? m->Vf.a[i] : C<double>(0.0));
          put(Vo[e] + 2 * (16 * k + 8 + i), m ? m->Vb.a[i] : C<double>(0.0));
        }
      if (Qo[e])
        for (int i = 0; i < 2; ++i) {
          put(Qo[e] + 2 * (4 * k + i), m ? m->Qf[i] : C<double>(0.0));
          put(Qo[e] + 2 * (4 * k + 2 + i), m ? m->Qb[i] : C<double>(0.0));
        }
    }
  }
  return 0;
}

// Batched _cavity_response at nodes (kx, ky) (kx = k_rho cos(phi), ky = k_rho
// sin(phi) as Modes_Stack(stack, omega, k_rho, phi)); s 1-based.
int ora_cavity_response(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                        double omega, int s, const double *kx, const double *ky, int64_t nn, double *out,
                        int nthreads) {
  if (s < 1 || s > n - 1) return 1;
  auto st = make_stack<double>(n, eps, mu, d, ek, mk);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t k = 0; k < nn; ++k) out[k] = cavity_response<double>(st, omega, s, kx[k], ky[k]);
  return 0;
}

// Type-1 nonuniform DFT in 2D, summed directly (the definition FINUFFT's
// type 1 with iflag +1 approximates; the 2D lateral sum of closed_form_G,
// mod_ClosedFormField.jl:743-780): for each listed mode (k1, k2),
//   f = sum_j c_j exp(+i (k1 x_j + k2 y_j)).
int ora_nudft2d1_modes(int64_t M, const double *x, const double *y, const double *c, int64_t nmodes, const int *k1,
                       const int *k2, double *f, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (int64_t m = 0; m < nmodes; ++m) {
    double sr = 0.0, si = 0.0;
    for (int64_t j = 0; j < M; ++j) {
      const double ph = k1[m] * x[j] + k2[m] * y[j];
      const double cr = std::cos(ph), ci = std::sin(ph);
      sr += c[2 * j] * cr - c[2 * j + 1] * ci;
      si += c[2 * j] * ci + c[2 * j + 1] * cr;
    }
    f[2 * m] = sr;
    f[2 * m + 1] = si;
  }
  return 0;
}

// The z coupling of apply_operator at every padded spectral node
// (_zmul_add!, _dense_add!, _zfold, _zexpand_add!, mod_GreenOperator.jl:604-662,
// and their use at :690-707), restated as direct sums over plane pairs (the
// reference forms the Toeplitz / Hankel sums as z convolutions by FFT; the sums
// are the same):
//   structured part: E[j] += sum_i (T[j - i + ns - 1] + H[i + j]) J[i]  (nT, nH > 0);
//   dense part:      E[j] += sum_i D[i + ns j] J[i]                     (Ps == NULL);
//   skeleton part:   Js[a] = sum_i Ps[i, a] J[i], Es[b] = sum_a D[a + rs b] Js[a],
//                    E[j] += sum_b Po[j, b] Es[b].
// Each product is the 3x3 plane [c][q] at the node times the 3-vector J[.][q].
// Layouts (complex, interleaved): planes [slot][c][q][nxy], J [ns][3][nxy],
// E [no][3][nxy] (accumulated), Ps [ns][rs], Po [no][ro] row-major.  OpenMP over
// chunks of nodes.
int ora_op_zmix(int64_t nxy, int ns, const double *J, int no, double *E, int nT, const double *T, int nH,
                const double *H, int rs, int ro, const double *D, const double *Ps, const double *Po, int nthreads) {
  using C = std::complex<double>;
  const C *Jc = (const C *)J, *Tc = (const C *)T, *Hc = (const C *)H, *Dc = (const C *)D, *Psc = (const C *)Ps,
          *Poc = (const C *)Po;
  C *Ec = (C *)E;
  const int64_t CH = 128;
  const int64_t nch = (nxy + CH - 1) / CH;
  const bool skel = Dc && Psc;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int64_t ch = 0; ch < nch; ++ch) {
    const int64_t x0 = ch * CH, nx = std::min<int64_t>(CH, nxy - x0);
    auto plane = [&](const C *P, int slot, int c, int q) { return P + ((int64_t)(slot * 9 + c * 3 + q)) * nxy + x0; };
    auto add = [&](C *out, const C *K, const C *v) {
      for (int c = 0; c < 3; ++c)
        for (int q = 0; q < 3; ++q) {
          const C *k = K + (int64_t)(c * 3 + q) * nxy;
          const C *x = v + (int64_t)q * nxy;
          C *o = out + (int64_t)c * nxy;
          for (int64_t t = 0; t < nx; ++t) o[t] += k[t] * x[t];
        }
    };
    if (nT > 0 || nH > 0)
      for (int j = 0; j < no; ++j)
        for (int i = 0; i < ns; ++i) {
          C *out = Ec + (int64_t)j * 3 * nxy + x0;
          const C *v = Jc + (int64_t)i * 3 * nxy + x0;
          if (nT > 0) add(out, plane(Tc, j - i + ns - 1, 0, 0), v);
          if (nH > 0) add(out, plane(Hc, i + j, 0, 0), v);
        }
    if (Dc && !skel)
      for (int j = 0; j < no; ++j)
        for (int i = 0; i < ns; ++i)
          add(Ec + (int64_t)j * 3 * nxy + x0, plane(Dc, i + ns * j, 0, 0), Jc + (int64_t)i * 3 * nxy + x0);
    if (skel) {
      // the folded sources and skeleton outputs of this chunk (stride CH)
      std::vector<C> Js((size_t)std::max(rs, 1) * 3 * CH), Es((size_t)std::max(ro, 1) * 3 * CH);
      for (int a = 0; a < rs; ++a)
        for (int q = 0; q < 3; ++q) {
          C *dst = Js.data() + ((size_t)a * 3 + q) * CH;
          for (int64_t t = 0; t < nx; ++t) dst[t] = 0.0;
          for (int i = 0; i < ns; ++i) {
            const C w = Psc[(int64_t)i * rs + a];
            const C *src = Jc + ((int64_t)i * 3 + q) * nxy + x0;
            for (int64_t t = 0; t < nx; ++t) dst[t] += w * src[t];
          }
        }
      std::fill(Es.begin(), Es.end(), C(0.0));
      for (int b = 0; b < ro; ++b)
        for (int a = 0; a < rs; ++a)
          for (int c = 0; c < 3; ++c)
            for (int q = 0; q < 3; ++q) {
              const C *k = plane(Dc, a + rs * b, c, q);
              const C *x = Js.data() + ((size_t)a * 3 + q) * CH;
              C *o = Es.data() + ((size_t)b * 3 + c) * CH;
              for (int64_t t = 0; t < nx; ++t) o[t] += k[t] * x[t];
            }
      for (int j = 0; j < no; ++j)
        for (int b = 0; b < ro; ++b) {
          const C w = Poc[(int64_t)j * ro + b];
          for (int c = 0; c < 3; ++c) {
            const C *src = Es.data() + ((size_t)b * 3 + c) * CH;
            C *o = Ec + ((int64_t)j * 3 + c) * nxy + x0;
            for (int64_t t = 0; t < nx; ++t) o[t] += w * src[t];
          }
        }
    }
  }
  return 0;
}

// Op count of one node solve + one spectral dyadic (reference formulation).
// out: add, mul, div, sqrt, trans.
int ora_count_node_ops(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                       double omega, double kx, double ky, int src, int obs_layer, double zsrc, double zobs,
                       uint64_t *out) {
  auto st = make_stack<CR>(n, eps, mu, d, ek, mk);
  std::vector<int> ob(1, obs_layer - 1);
  g_ops = Ops();
  NodeOut<CR> nd;
  node_solve(st, CR(omega), CR(kx), CR(ky), src - 1, ob, nd);
  Ops solve = g_ops;
  cx<CR> gg[9];
  node_ghat(nd, 0, obs_layer == src, src - 1, d[src - 1], d[obs_layer - 1], zsrc, zobs, gg);
  out[0] = solve.add; out[1] = solve.mul; out[2] = solve.div; out[3] = solve.sqrt; out[4] = solve.trans;
  out[5] = g_ops.add - solve.add; out[6] = g_ops.mul - solve.mul; out[7] = g_ops.div - solve.div;
  out[8] = g_ops.sqrt - solve.sqrt; out[9] = g_ops.trans - solve.trans;
  return 0;
}

// Op count of Modes_Stack + the total S (SLeft[N]) of S_all_Matrix for one
// node, in the reference's formulation (general eigen-solve per non-closed-form
// layer, full 4x4 star products).  out: add, mul, div, sqrt, trans.
int ora_count_stack_ops(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                        double omega, double kx, double ky, uint64_t *out) {
  auto st = make_stack<CR>(n, eps, mu, d, ek, mk);
  g_ops = Ops();
  std::vector<Modes<CR>> M;
  modes_stack(st, CR(omega), CR(kx), CR(ky), M);
  // S_all_ingredients + the SLeft prefix scan of S_all_Matrix up to SLeft[N]
  // (the suffix scan is not needed for the total S)
  Mat<CR, 4, 4> S0 = eye<CR, 4>(), St;
  {
    cx<CR> dg[4];
    S_prop(M[0], st[0].d, dg);
    S0 = redheffer(S0, diag4(dg));
  }
  for (int j = 1; j < n; ++j) {
    Mat<CR, 4, 4> Si = S_interface(M[j - 1].Vf, M[j - 1].Vb, M[j].Vf, M[j].Vb);
    St = redheffer(S0, Si);
    if (j < n - 1) {
      cx<CR> dg[4];
      S_prop(M[j], st[j].d, dg);
      S0 = redheffer(St, diag4(dg));
    }
  }
  out[0] = g_ops.add; out[1] = g_ops.mul; out[2] = g_ops.div; out[3] = g_ops.sqrt; out[4] = g_ops.trans;
  return 0;
}

// Op count of one _cavity_response evaluation (reference formulation:
// Modes_Stack, S_all_ingredients, both scans of S_all_Matrix, the 2x2 parts).
// out: add, mul, div, sqrt, trans.
int ora_count_response_ops(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                           double omega, int s, double kx, double ky, uint64_t *out) {
  if (s < 1 || s > n - 1) return 1;
  auto st = make_stack<CR>(n, eps, mu, d, ek, mk);
  g_ops = Ops();
  cavity_response<CR>(st, CR(omega), s, CR(kx), CR(ky));
  out[0] = g_ops.add; out[1] = g_ops.mul; out[2] = g_ops.div; out[3] = g_ops.sqrt; out[4] = g_ops.trans;
  return 0;
}

// Modal_amp for one ModalSeed (test-only reference route).  Ef, Eb: n*2 complex.
int ora_modal_amp(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                  double omega, double kx, double ky, int layer, double z, const double *sif, const double *sib,
                  double *Ef, double *Eb) {
  auto st = make_stack<double>(n, eps, mu, d, ek, mk);
  std::vector<Modes<double>> M;
  if (!modes_stack(st, omega, kx, ky, M)) return 1;
  Smat<double> S;
  S_all(st, M, S);
  cx<double> a[2] = {get<double>(sif), get<double>(sif + 2)}, b[2] = {get<double>(sib), get<double>(sib + 2)};
  std::vector<Mat<double, 2, 1>> ef, eb;
  modal_amp(st, M, S, layer - 1, z, a, b, ef, eb);
  for (int l = 0; l < n; ++l) {
    put(Ef + 4 * l, ef[l](0, 0)); put(Ef + 4 * l + 2, ef[l](1, 0));
    put(Eb + 4 * l, eb[l](0, 0)); put(Eb + 4 * l + 2, eb[l](1, 0));
  }
  return 0;
}

// Modal_amp_profile (mod_FieldSolver.jl:375-406) at local depths zs (one seed).
int ora_modal_amp_profile(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                          double omega, double kx, double ky, int layer, double z, const double *sif,
                          const double *sib, const double *zs, int nz, double *Ef, double *Eb) {
  auto st = make_stack<double>(n, eps, mu, d, ek, mk);
  std::vector<Modes<double>> M;
  if (!modes_stack(st, omega, kx, ky, M)) return 1;
  Smat<double> S;
  S_all(st, M, S);
  int L = layer - 1;
  if (L >= n - 1) return 2;
  cx<double> a[2] = {get<double>(sif), get<double>(sif + 2)}, b[2] = {get<double>(sib), get<double>(sib + 2)};
  std::vector<Mat<double, 2, 1>> ef, eb;
  Mat<double, 2, 1> Fp, Fm;
  Cavity<double> cv;
  modal_amp(st, M, S, L, z, a, b, ef, eb, &Fp, &Fm, &cv);
  double dl = st[L].d;
  double zpos = L == 0 ? dl - z : z;
  Mat<double, 2, 1> Ej0 = mul(cv.RL, Fm), Ebd = mul(cv.RR, Fp);
  for (int i = 0; i < nz; ++i) {
    Mat<double, 2, 1> e1 = Ej0, e2 = Ebd;
    if (zpos < zs[i]) {
      cx<double> ph[2];
      sprop2(-zpos, M[L].Qf, ph);
      e1(0, 0) = e1(0, 0) + a[0] * ph[0]; e1(1, 0) = e1(1, 0) + a[1] * ph[1];
    }
    if (zpos > zs[i]) {
      cx<double> ph[2];
      sprop2(dl - zpos, M[L].Qb, ph);
      e2(0, 0) = e2(0, 0) + b[0] * ph[0]; e2(1, 0) = e2(1, 0) + b[1] * ph[1];
    }
    e1 = flush(e1); e2 = flush(e2);
    put(Ef + 4 * i, e1(0, 0)); put(Ef + 4 * i + 2, e1(1, 0));
    put(Eb + 4 * i, e2(0, 0)); put(Eb + 4 * i + 2, e2(1, 0));
  }
  return 0;
}

// Cavity edge seeds (S+, S-) of one ModalSeed (_seed_from_sources).
int ora_seed_from_source(int n, const double *eps, const double *mu, const double *d, const int *ek, const int *mk,
                         double omega, double kx, double ky, int layer, double z, const double *sif,
                         const double *sib, double *u4) {
  auto st = make_stack<double>(n, eps, mu, d, ek, mk);
  std::vector<Modes<double>> M;
  if (!modes_stack(st, omega, kx, ky, M)) return 1;
  int L = layer - 1;
  double dl = st[L].d;
  double zpos = L == 0 ? dl - z : z;
  cx<double> sp[2] = {C<double>(0), C<double>(0)}, sm[2];
  if (L < n - 1) {
    cx<double> ph[2];
    sprop2(dl - zpos, M[L].Qf, ph);
    sp[0] = get<double>(sif) * ph[0]; sp[1] = get<double>(sif + 2) * ph[1];
  }
  cx<double> pb[2];
  sprop2(-zpos, M[L].Qb, pb);
  sm[0] = get<double>(sib) * pb[0]; sm[1] = get<double>(sib + 2) * pb[1];
  put(u4, sp[0]); put(u4 + 2, sp[1]); put(u4 + 4, sm[0]); put(u4 + 6, sm[1]);
  return 0;
}

// 1D Jacobi-Anger synthesis, closed_form_G(::CavityTransferCache1D)
// (mod_ClosedFormField.jl:435-487, 344-390).  Nodes kx (ky = 0) with weights
// wt and the oracle cache of those nodes.  G out: (nx, ny, nz, 3, 3) column-
// major complex.  with_direct = 0 keeps the reflected part only
// (terms=:scattered, mod_ClosedFormField.jl:143-160).
int ora_closed_form_G_1d(int n, const double *d, double omega, int src, const int *obs, int nobs, int obs_layer,
                         const double *kx, const double *wt, const double *rec, const int *valid, const int *Tvalid,
                         int64_t nk, double zsrc_local, const double *zo, int nz, const double *xs, int nx,
                         const double *ys, int ny, int with_direct, double *G, int nthreads) {
  (void)omega; (void)n;
  typedef cx<double> c;
  int rl = rec_len(nobs);
  int oi = -1;
  for (int i = 0; i < nobs; ++i)
    if (obs[i] == obs_layer) oi = i;
  if (oi < 0) return 1;
  bool same = obs_layer == src;
  double d_src = d[src - 1], d_obs = d[obs_layer - 1];
  double zpos = src == 1 ? d_src - zsrc_local : zsrc_local;
  std::vector<NodeOut<double>> nodes(nk);
  for (int64_t k = 0; k < nk; ++k) read_rec(nodes[k], nobs, rec + 2 * rl * k, valid[k], Tvalid + nobs * k);
  // amplitude_at (reflected part), per node/mode/moment
  std::vector<c> Afr(nk * 6), Abr(nk * 6);
  for (int64_t k = 0; k < nk; ++k) {
    NodeOut<double> &nd = nodes[k];
    for (int i = 0; i < 6; ++i) Afr[6 * k + i] = Abr[6 * k + i] = C<double>(0);
    if (!(nd.valid && nd.Tvalid[oi])) continue;
    c Pf[2], Pb[2];
    for (int m = 0; m < 2; ++m) {
      Pf[m] = cexp(times_i(nd.Qf[m]) * C<double>(d_src - zpos));
      Pb[m] = cexp(-times_i(nd.Qb[m]) * C<double>(zpos));
    }
    for (int q = 0; q < 3; ++q) {
      c u[4] = {nd.sf(0, q) * Pf[0], nd.sf(1, q) * Pf[1], nd.sb(0, q) * Pb[0], nd.sb(1, q) * Pb[1]};
      c v[4];
      for (int i = 0; i < 4; ++i) {
        c s = nd.T[oi](i, 0) * u[0];
        for (int j = 1; j < 4; ++j) s = s + nd.T[oi](i, j) * u[j];
        v[i] = s;
      }
      Afr[6 * k + 0 + 2 * q] = v[0]; Afr[6 * k + 1 + 2 * q] = v[1];
      Abr[6 * k + 0 + 2 * q] = v[2]; Abr[6 * k + 1 + 2 * q] = v[3];
    }
  }
  // distinct radii
  std::vector<double> rho;
  std::vector<int> rmap((size_t)nx * ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double r = std::hypot(xs[i], ys[j]);
      int idx = -1;
      for (size_t t = 0; t < rho.size(); ++t)
        if (rho[t] == r) { idx = (int)t; break; }
      if (idx < 0) { idx = (int)rho.size(); rho.push_back(r); }
      rmap[i + (size_t)nx * j] = idx;
    }
  int nr = (int)rho.size();
  std::vector<double> J0((size_t)nk * nr), J1((size_t)nk * nr), J2((size_t)nk * nr);
  for (int r = 0; r < nr; ++r)
    for (int64_t k = 0; k < nk; ++k) {
      double x = kx[k] * rho[r];
      J0[k + nk * r] = std::cyl_bessel_j(0.0, std::fabs(x)) ;
      double j1 = std::cyl_bessel_j(1.0, std::fabs(x));
      J1[k + nk * r] = x < 0 ? -j1 : j1;
      J2[k + nk * r] = std::cyl_bessel_j(2.0, std::fabs(x));
    }
  std::vector<c> Rm((size_t)9 * nr);
  for (int zi = 0; zi < nz; ++zi) {
    // weighted amplitudes (and phases) for this row
    std::vector<c> Ax(nk * 6), Axb(nk * 6), Pf(nk * 2), Pb(nk * 2);
    for (int64_t k = 0; k < nk; ++k) {
      NodeOut<double> &nd = nodes[k];
      const c *B = &nd.basis[16 * oi];
      c W = C<double>(0.0, wt[k] / (2.0 * M_PI));
      c Afd[6], Abd[6];
      for (int i = 0; i < 6; ++i) Afd[i] = Abd[i] = C<double>(0);
      if (same) {
        if (nd.valid && with_direct) {
          double zeta = zo[zi] - zpos;
          for (int m = 0; m < 2; ++m) {
            if (zeta > 0) {
              c e = phase(nd.Qf[m], zeta);
              for (int q = 0; q < 3; ++q) Afd[m + 2 * q] = nd.sf(m, q) * e;
            } else if (zeta < 0) {
              c e = phase(nd.Qb[m], zeta);
              for (int q = 0; q < 3; ++q) Abd[m + 2 * q] = nd.sb(m, q) * e;
            }
          }
        }
        for (int m = 0; m < 2; ++m) {
          c pf = phase(B[12 + m], zo[zi]), pb = phase(B[14 + m], zo[zi] - d_obs);
          for (int q = 0; q < 3; ++q) {
            c af = Afd[m + 2 * q] + Afr[6 * k + m + 2 * q] * pf;
            c ab = Abd[m + 2 * q] + Abr[6 * k + m + 2 * q] * pb;
            Ax[6 * k + m + 2 * q] = finite_or_zero(W * af);
            Axb[6 * k + m + 2 * q] = finite_or_zero(W * ab);
          }
          Pf[2 * k + m] = C<double>(1.0); Pb[2 * k + m] = C<double>(1.0);
        }
      } else {
        for (int m = 0; m < 2; ++m) {
          for (int q = 0; q < 3; ++q) {
            Ax[6 * k + m + 2 * q] = finite_or_zero(W * Afr[6 * k + m + 2 * q]);
            Axb[6 * k + m + 2 * q] = finite_or_zero(W * Abr[6 * k + m + 2 * q]);
          }
          Pf[2 * k + m] = phase(B[12 + m], zo[zi]);
          Pb[2 * k + m] = phase(B[14 + m], zo[zi] - d_obs);
        }
      }
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int r = 0; r < nr; ++r) {
      double rh = rho[r];
      double cpsi = rh == 0.0 ? 0.0 : rh / rh;
      for (int q = 0; q < 3; ++q) {
        c px = C<double>(q == 0), py = C<double>(q == 1), pz = C<double>(q == 2);
        c vx = C<double>(0), vy = C<double>(0), vz = C<double>(0);
        c ihalf = C<double>(0.0, 0.5), mi = C<double>(0.0, -1.0);
        for (int64_t k = 0; k < nk; ++k) {
          double j0 = J0[k + nk * r], j1 = J1[k + nk * r], j2 = J2[k + nk * r];
          if (!(std::isfinite(j0) && std::isfinite(j1) && std::isfinite(j2))) continue;
          const c *B = &nodes[k].basis[16 * oi];
          for (int m = 0; m < 2; ++m) {
            c exf = B[0 + m], eyf = B[2 + m], ezf = B[4 + m], exb = B[6 + m], eyb = B[8 + m], ezb = B[10 + m];
            c axf = Ax[6 * k + m + 0], ayf = Ax[6 * k + m + 2], azf = Ax[6 * k + m + 4];
            c Af = px * axf + py * ayf, Bf = px * ayf - py * axf, Cf = pz * azf;
            c J20 = C<double>(j2 - j0), J02 = C<double>(j0 + j2), J1c = C<double>(j1 * cpsi), J0c = C<double>(j0);
            c Ixf = ihalf * (J20 * Af * exf - J02 * Bf * eyf) + J1c * Cf * exf;
            c Iyf = ihalf * (J20 * Af * eyf + J02 * Bf * exf) + J1c * Cf * eyf;
            c Izf = J1c * Af * ezf + (mi * J0c) * Cf * ezf;
            c axb = Axb[6 * k + m + 0], ayb = Axb[6 * k + m + 2], azb = Axb[6 * k + m + 4];
            c Ab = px * axb + py * ayb, Bb = px * ayb - py * axb, Cb = pz * azb;
            c Ixb = ihalf * (J20 * Ab * exb - J02 * Bb * eyb) + J1c * Cb * exb;
            c Iyb = ihalf * (J20 * Ab * eyb + J02 * Bb * exb) + J1c * Cb * eyb;
            c Izb = J1c * Ab * ezb + (mi * J0c) * Cb * ezb;
            c pf = Pf[2 * k + m], pb = Pb[2 * k + m];
            c kk = C<double>(kx[k]);
            vx = vx + (Ixf * pf + Ixb * pb) * kk;
            vy = vy + (Iyf * pf + Iyb * pb) * kk;
            vz = vz + (Izf * pf + Izb * pb) * kk;
          }
        }
        Rm[9 * r + 0 + 3 * q] = vx; Rm[9 * r + 1 + 3 * q] = vy; Rm[9 * r + 2 + 3 * q] = vz;
      }
    }
    // per-pixel rotation Rot * R(rho) * Rot^T
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        double r = std::hypot(xs[i], ys[j]);
        double cp = r == 0.0 ? 1.0 : xs[i] / r, sp = r == 0.0 ? 0.0 : ys[j] / r;
        double Rot[9] = {cp, sp, 0, -sp, cp, 0, 0, 0, 1};  // col-major
        const c *Mr = &Rm[9 * rmap[i + (size_t)nx * j]];
        for (int q = 0; q < 3; ++q)
          for (int cc = 0; cc < 3; ++cc) {
            c s = C<double>(0);
            for (int a = 0; a < 3; ++a)
              for (int b = 0; b < 3; ++b)
                s = s + C<double>(Rot[cc + 3 * a] * Rot[q + 3 * b]) * Mr[a + 3 * b];
            size_t off = i + (size_t)nx * (j + (size_t)ny * (zi + (size_t)nz * (cc + 3 * (size_t)q)));
            put(G + 2 * off, s);
          }
      }
  }
  return 0;
}

}  // extern "C"
