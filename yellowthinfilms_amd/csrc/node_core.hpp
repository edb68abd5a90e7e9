// node_core.hpp -- the batched Green k-node solve on gfx950, FP64 (device
// code shared by node_top.hip and node_general.hip).
//
// One lane solves one in-plane wavevector node (kx, ky) of a planar stack:
//   Berreman 4x4 modes of every layer (closed form for isotropic/uniaxial
//   layers, general complex eigen-decomposition otherwise), forward/backward
//   classification and continuity matching, interface S-matrices, the
//   Redheffer cascade, the cavity resummation of the source layer, the walk to
//   the observation layer (T^o), the point-source seed projection and the
//   observation layer's mode-field basis; optionally the spectral dyadic
//   g~(k; z, z') for a list of (source depth, observation depth) pairs.
//
// Reference behaviour followed (HortenseBeaudoin/YellowThinFilms):
//   modes            mod_ModeSolver.jl:122-193, 203-265, 274-349, 358-392
//   S-matrices       mod_ScatteringMatrix.jl:27-154
//   seed projection  mod_SourceProj.jl:34-60
//   cavity, walk, T  mod_FieldSolver.jl:24-53, 72-152, 191-341
//   node + basis     mod_CavityTransfer.jl:87-130, 224-251, 339-359
//   spectral dyadic  mod_ClosedFormField.jl:59-69, 539-602, 616-659
//
// GPU-specific restructuring (same mathematics):
//   * Single forward pass over the layers.  SRight[k] (the reference's
//     backward suffix scan) is accumulated left to right as
//     eff_k * eff_{k+1} * ... with eff_k = Sprop[k] * Sint[k] ("*" the Redheffer
//     star product, which is associative), so no per-layer S-matrix storage
//     is needed; only the suffixes a node actually uses are kept.
//   * The source and observation layers' modes are re-solved after the pass
//     (with the recorded matching swaps), so their 4x4 mode matrices are not
//     held in registers through the cascade.
//   * The Redheffer product with a diagonal propagation matrix is specialised
//     (its 2x2 inverses are exact identities).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ytf_device.hpp"
#include "ytf_kernels.h"

namespace ytf {

// ---------------------------------------------------------------------------
// Berreman matrix -- mod_ModeSolver.jl:122-193
// ---------------------------------------------------------------------------
__device__ __forceinline__ void berreman(const LayerDev &L, double w, double kx, double ky, cd *B) {
  const cd e11 = L.eps[0], e21 = L.eps[1], e31 = L.eps[2], e12 = L.eps[3], e22 = L.eps[4], e32 = L.eps[5];
  const cd e13 = L.eps[6], e23 = L.eps[7], e33 = L.eps[8];
  const cd m11 = L.mu[0], m21 = L.mu[1], m31 = L.mu[2], m12 = L.mu[3], m22 = L.mu[4], m32 = L.mu[5];
  const cd m13 = L.mu[6], m23 = L.mu[7], m33 = L.mu[8];
  const cd k1 = mk(kx), k2 = mk(ky), om = mk(w), one = mk(1.0);
  const cd denE = one / e33, denM = one / m33, iw = one / om;
  const cd Se11 = e11 - e13 * denE * e31, Se12 = e12 - e13 * denE * e32;
  const cd Se21 = e21 - e23 * denE * e31, Se22 = e22 - e23 * denE * e32;
  const cd Sm11 = m11 - m13 * denM * m31, Sm12 = m12 - m13 * denM * m32;
  const cd Sm21 = m21 - m23 * denM * m31, Sm22 = m22 - m23 * denM * m32;
  const cd rE1 = denE * e31, rE2 = denE * e32, sE1 = e13 * denE, sE2 = e23 * denE;
  const cd rM1 = denM * m31, rM2 = denM * m32, sM1 = m13 * denM, sM2 = m23 * denM;
  const cd t21 = -conj(k2), t22 = conj(k1);
  const cd t31 = iw * k1, t32 = iw * k2;
  const cd RsM1 = sM2, RsM2 = -sM1, RsE1 = sE2, RsE2 = -sE1;
  B[0] = RsM1 * t21 - k1 * rE1;        // B11_11
  B[4] = RsM1 * t22 - k1 * rE2;        // B11_12
  B[1] = RsM2 * t21 - k2 * rE1;        // B11_21
  B[5] = RsM2 * t22 - k2 * rE2;        // B11_22
  B[8] = om * Sm21 - t31 * denE * t21;   // B12_11
  B[12] = om * Sm22 - t31 * denE * t22;  // B12_12
  B[9] = -om * Sm11 - t32 * denE * t21;  // B12_21
  B[13] = -om * Sm12 - t32 * denE * t22; // B12_22
  B[2] = t31 * denM * t21 - om * Se21;   // B21_11
  B[6] = t31 * denM * t22 - om * Se22;   // B21_12
  B[3] = t32 * denM * t21 + om * Se11;   // B21_21
  B[7] = t32 * denM * t22 + om * Se12;   // B21_22
  B[10] = RsE1 * t21 - k1 * rM1;       // B22_11
  B[14] = RsE1 * t22 - k1 * rM2;       // B22_12
  B[11] = RsE2 * t21 - k2 * rM1;       // B22_21
  B[15] = RsE2 * t22 - k2 * rM2;       // B22_22
}

// ---------------------------------------------------------------------------
// General complex 4x4 eigen-decomposition (stands in for LAPACK geevx):
// diagonal balancing, Householder Hessenberg, shifted QR to Schur form with
// accumulated Z, triangular back-substitution, back-transform, zgeev-style
// normalisation (unit 2-norm, largest component real positive), then the
// (Re, Im) ascending order of Julia's eigen.  Returns false on no convergence.
// ---------------------------------------------------------------------------
template <int IHI>
__device__ __forceinline__ bool qr_deflate(cd *A, cd *Z) {
  const double ulp = 2.220446049250313e-16;
#define AA(i, j) A[(i) + 4 * (j)]
#define ZZ(i, j) Z[(i) + 4 * (j)]
  for (int it = 0; it < 60; ++it) {
    double hk = cabs1(AA(IHI, IHI - 1));
    double tst = cabs1(AA(IHI - 1, IHI - 1)) + cabs1(AA(IHI, IHI));
    if (tst == 0.0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) tst += cabs1(AA(j, IHI));
    }
    if (hk <= ulp * tst) {
      AA(IHI, IHI - 1) = mk(0.0);
      return true;
    }
    cd mu;
    if (it % 11 == 10) {
      mu = AA(IHI, IHI) + mk(0.75 * cabs(AA(IHI, IHI - 1)));
    } else {
      const cd a = AA(IHI - 1, IHI - 1), b = AA(IHI - 1, IHI), c = AA(IHI, IHI - 1), d = AA(IHI, IHI);
      const cd tr2 = rmul(0.5, a + d), dd = rmul(0.5, a - d);
      const cd disc = csqrt(dd * dd + b * c);
      const cd l1 = tr2 + disc, l2 = tr2 - disc;
      mu = sel(abs2(l1 - d) <= abs2(l2 - d), l1, l2);
    }
#pragma unroll
    for (int k = 0; k <= IHI; ++k) AA(k, k) = AA(k, k) - mu;
    double cs[IHI];
    cd sn[IHI];
#pragma unroll
    for (int k = 0; k < IHI; ++k) {
      const cd a = AA(k, k), b = AA(k + 1, k);
      const double na = cabs(a), nb = cabs(b);
      const double nrm = hypot(na, nb);
      double c = 1.0;
      cd s = mk(0.0);
      if (nrm != 0.0) {
        if (na == 0.0) {
          c = 0.0;
          s = mk(1.0);
        } else {
          c = na / nrm;
          s = rmul(1.0 / (na * nrm), a * conj(b));
        }
      }
      cs[k] = c;
      sn[k] = s;
#pragma unroll
      for (int j = k; j < 4; ++j) {
        const cd x = AA(k, j), y = AA(k + 1, j);
        AA(k, j) = rmul(c, x) + s * y;
        AA(k + 1, j) = rmul(c, y) - conj(s) * x;
      }
    }
#pragma unroll
    for (int k = 0; k < IHI; ++k) {
      const double c = cs[k];
      const cd s = sn[k];
#pragma unroll
      for (int i = 0; i <= (k + 1 < IHI ? k + 1 : IHI); ++i) {
        const cd x = AA(i, k), y = AA(i, k + 1);
        AA(i, k) = rmul(c, x) + y * conj(s);
        AA(i, k + 1) = rmul(c, y) - x * s;
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const cd x = ZZ(i, k), y = ZZ(i, k + 1);
        ZZ(i, k) = rmul(c, x) + y * conj(s);
        ZZ(i, k + 1) = rmul(c, y) - x * s;
      }
    }
#pragma unroll
    for (int k = 0; k <= IHI; ++k) AA(k, k) = AA(k, k) + mu;
  }
  return false;
#undef AA
#undef ZZ
}

__device__ __forceinline__ void cswap_col(bool c, cd *Q, cd *V, int a, int b) {
  const cd qa = Q[a], qb = Q[b];
  Q[a] = sel(c, qb, qa);
  Q[b] = sel(c, qa, qb);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const cd va = V[i + 4 * a], vb = V[i + 4 * b];
    V[i + 4 * a] = sel(c, vb, va);
    V[i + 4 * b] = sel(c, va, vb);
  }
}
__device__ __forceinline__ bool lessRI(cd a, cd b) { return a.r < b.r || (a.r == b.r && a.i < b.i); }

__device__ __forceinline__ bool eigen4(cd *A, cd *Q, cd *V) {
  const double ulp = 2.220446049250313e-16;
#define AA(i, j) A[(i) + 4 * (j)]
  // balance (EISPACK balanc, radix 2, no permutation)
  double D[4] = {1.0, 1.0, 1.0, 1.0};
  for (int sweep = 0; sweep < 32; ++sweep) {
    bool conv = true;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double c = 0.0, r = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j == i) continue;
        c += cabs1(AA(j, i));
        r += cabs1(AA(i, j));
      }
      if (c == 0.0 || r == 0.0) continue;
      double g = r * 0.5, f = 1.0;
      const double s = c + r;
      while (c < g) { f *= 2.0; c *= 4.0; }
      g = r * 2.0;
      while (c >= g) { f *= 0.5; c *= 0.25; }
      if ((c + r) / f < 0.95 * s) {
        conv = false;
        D[i] *= f;
        const double fi = 1.0 / f;
#pragma unroll
        for (int j = 0; j < 4; ++j) AA(i, j) = rmul(fi, AA(i, j));
#pragma unroll
        for (int j = 0; j < 4; ++j) AA(j, i) = rmul(f, AA(j, i));
      }
    }
    if (conv) break;
  }
  // Householder reduction to Hessenberg form, accumulating Z
  cd Z[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) Z[i] = mk((i % 5 == 0) ? 1.0 : 0.0);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    double xn2 = 0.0;
#pragma unroll
    for (int i = k + 1; i < 4; ++i) xn2 += abs2(AA(i, k));
    if (xn2 != 0.0) {
      const double xn = sqrt(xn2);
      const cd x0 = AA(k + 1, k);
      const double ax0 = cabs(x0);
      const cd ph = ax0 == 0.0 ? mk(1.0) : rmul(1.0 / ax0, x0);
      const cd alpha = rmul(-xn, ph);
      cd v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = (i < k + 1) ? mk(0.0) : AA(i, k);
      v[k + 1] = x0 - alpha;
      double vn = 0.0;
#pragma unroll
      for (int i = k + 1; i < 4; ++i) vn += abs2(v[i]);
      if (vn != 0.0) {
        const double tau = 2.0 / vn;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          cd s = mk(0.0);
#pragma unroll
          for (int i = k + 1; i < 4; ++i) s = fmac_cj(s, v[i], AA(i, j));
          s = rmul(tau, s);
#pragma unroll
          for (int i = k + 1; i < 4; ++i) AA(i, j) = AA(i, j) - v[i] * s;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          cd s = mk(0.0), sz = mk(0.0);
#pragma unroll
          for (int j = k + 1; j < 4; ++j) {
            s = fmac(s, AA(i, j), v[j]);
            sz = fmac(sz, Z[i + 4 * j], v[j]);
          }
          s = rmul(tau, s);
          sz = rmul(tau, sz);
#pragma unroll
          for (int j = k + 1; j < 4; ++j) {
            AA(i, j) = AA(i, j) - s * conj(v[j]);
            Z[i + 4 * j] = Z[i + 4 * j] - sz * conj(v[j]);
          }
        }
#pragma unroll
        for (int i = k + 2; i < 4; ++i) AA(i, k) = mk(0.0);
      }
    }
  }
  bool ok = qr_deflate<3>(A, Z);
  ok = ok && qr_deflate<2>(A, Z);
  ok = ok && qr_deflate<1>(A, Z);
  // eigenvectors of the upper-triangular Schur factor, back-transformed
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const cd lam = AA(k, k);
    cd x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = mk(i == k ? 1.0 : 0.0);
    double smin = ulp * cabs(lam);
    smin = smin > 1e-290 ? smin : 1e-290;
#pragma unroll
    for (int j = k - 1; j >= 0; --j) {
      cd s = mk(0.0);
#pragma unroll
      for (int l = j + 1; l <= k; ++l) s = fmac(s, AA(j, l), x[l]);
      cd den = AA(j, j) - lam;
      if (cabs(den) < smin) den = mk(smin);
      x[j] = -(s / den);
    }
    double s2 = 0.0;
    cd col[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cd s = mk(0.0);
#pragma unroll
      for (int j = 0; j <= k; ++j) s = fmac(s, Z[i + 4 * j], x[j]);
      col[i] = rmul(D[i], s);
      s2 += abs2(col[i]);
    }
    const double inv = 1.0 / sqrt(s2);
    // zgeev normalisation: unit 2-norm, the largest component made real and
    // positive.  Near-ties (|v_i| equal to rounding, e.g. circular
    // polarisation at normal incidence) take the first such component so the
    // choice does not depend on rounding.
    double best = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      col[i] = rmul(inv, col[i]);
      best = fmax(best, abs2(col[i]));
    }
    const double thr = best * (1.0 - 1e-10);
    int imax = 3;
    cd p = col[3];
#pragma unroll
    for (int i = 2; i >= 0; --i) {
      const bool t = abs2(col[i]) >= thr;
      imax = t ? i : imax;
      p = sel(t, col[i], p);
    }
    const double ap = cabs(p);
    const cd rot = mk(p.r / ap, -(p.i / ap));
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const cd v = col[i] * rot;
      V[i + 4 * k] = cd{v.r, i == imax ? 0.0 : v.i};
    }
    Q[k] = lam;
  }
#undef AA
  // sort eigenpairs by (Re, Im): 4-element sorting network
  cswap_col(lessRI(Q[1], Q[0]), Q, V, 0, 1);
  cswap_col(lessRI(Q[3], Q[2]), Q, V, 2, 3);
  cswap_col(lessRI(Q[2], Q[0]), Q, V, 0, 2);
  cswap_col(lessRI(Q[3], Q[1]), Q, V, 1, 3);
  cswap_col(lessRI(Q[2], Q[1]), Q, V, 1, 2);
  return ok;
}

// Eigen-decomposition of a block off-diagonal Berreman matrix
// B = [0 P; R 0] (P = B12, R = B21): the form B takes whenever eps and mu have
// no xz/yz coupling (isotropic, uniaxial, diagonal biaxial, gyrotropic about z,
// any in-plane rotation of those).  Then B^2 = diag(P R, R P), so q^2 are the
// eigenvalues lam of the 2x2 M = P R and, for q = +-sqrt(lam) with M a = lam a,
// the eigenvector is v = [a; R a / q].  lam2 = det(P) det(R) / lam1 avoids the
// cancellation of the quadratic formula.  Vectors get the same normalisation
// and ordering as eigen4 (zgeev style, sorted by (Re, Im)).
__device__ __forceinline__ void normalise_col(cd *col) {
  double s2 = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) s2 += abs2(col[i]);
  const double inv = 1.0 / sqrt(s2);
  double best = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    col[i] = rmul(inv, col[i]);
    best = fmax(best, abs2(col[i]));
  }
  const double thr = best * (1.0 - 1e-10);
  int imax = 3;
  cd p = col[3];
#pragma unroll
  for (int i = 2; i >= 0; --i) {
    const bool t = abs2(col[i]) >= thr;
    imax = t ? i : imax;
    p = sel(t, col[i], p);
  }
  const double ap = cabs(p);
  const cd rot = mk(p.r / ap, -(p.i / ap));
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const cd v = col[i] * rot;
    col[i] = cd{v.r, i == imax ? 0.0 : v.i};
  }
}

__device__ __forceinline__ bool is_block_form(const LayerDev &L) {
  const int idx[4] = {2, 5, 6, 7};  // (3,1) (3,2) (1,3) (2,3), column-major
  bool z = true;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    z = z && L.eps[idx[t]].r == 0.0 && L.eps[idx[t]].i == 0.0;
    z = z && L.mu[idx[t]].r == 0.0 && L.mu[idx[t]].i == 0.0;
  }
  return z;
}

__device__ __forceinline__ bool eigen_block(const cd *B, cd *Q, cd *V) {
  // P = B[0:2, 2:4], R = B[2:4, 0:2] (column-major 4x4)
  const cd P11 = B[8], P21 = B[9], P12 = B[12], P22 = B[13];
  const cd R11 = B[2], R21 = B[3], R12 = B[6], R22 = B[7];
  const cd m11 = P11 * R11 + P12 * R21, m12 = P11 * R12 + P12 * R22;
  const cd m21 = P21 * R11 + P22 * R21, m22 = P21 * R12 + P22 * R22;
  const cd half_tr = rmul(0.5, m11 + m22), dd = rmul(0.5, m11 - m22);
  cd disc = csqrt(dd * dd + m12 * m21);
  // larger-magnitude root first
  const cd la = half_tr + disc, lb = half_tr - disc;
  const cd l1 = sel(abs2(la) >= abs2(lb), la, lb);
  const cd detM = (P11 * P22 - P12 * P21) * (R11 * R22 - R12 * R21);
  const cd l2 = (abs2(l1) > 0.0) ? detM / l1 : mk(0.0);
  const cd lam[2] = {l1, l2};
  cd a[2][2];
  const double scale = abs2(m11) + abs2(m12) + abs2(m21) + abs2(m22);
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    // (m12, lam - m11) and (lam - m22, m21) both solve (M - lam) a = 0; take the larger
    const cd u1a = m12, u1b = lam[e] - m11, u2a = lam[e] - m22, u2b = m21;
    const double n1 = abs2(u1a) + abs2(u1b), n2 = abs2(u2a) + abs2(u2b);
    const bool use1 = n1 >= n2;
    cd x = sel(use1, u1a, u2a), y = sel(use1, u1b, u2b);
    // M = lam I (degenerate, non-defective): the coordinate vectors
    const bool flat = fmax(n1, n2) <= 1e-28 * scale;
    x = sel(flat, mk(e == 0 ? 1.0 : 0.0), x);
    y = sel(flat, mk(e == 0 ? 0.0 : 1.0), y);
    a[e][0] = x;
    a[e][1] = y;
  }
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const cd q = csqrt(lam[e]);
    const cd Ra0 = R11 * a[e][0] + R12 * a[e][1], Ra1 = R21 * a[e][0] + R22 * a[e][1];
    const cd iq = cinv(q);
#pragma unroll
    for (int sgn = 0; sgn < 2; ++sgn) {
      const cd qs = sgn == 0 ? q : -q;
      const cd iqs = sgn == 0 ? iq : -iq;
      cd col[4] = {a[e][0], a[e][1], Ra0 * iqs, Ra1 * iqs};
      normalise_col(col);
      const int k = 2 * e + sgn;
      Q[k] = qs;
#pragma unroll
      for (int i = 0; i < 4; ++i) V[i + 4 * k] = col[i];
    }
  }
  cswap_col(lessRI(Q[1], Q[0]), Q, V, 0, 1);
  cswap_col(lessRI(Q[3], Q[2]), Q, V, 2, 3);
  cswap_col(lessRI(Q[2], Q[0]), Q, V, 0, 2);
  cswap_col(lessRI(Q[3], Q[1]), Q, V, 1, 3);
  cswap_col(lessRI(Q[2], Q[1]), Q, V, 1, 2);
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 4; ++i) ok = ok && isfin(Q[i]);
  return ok;
}

// Closed-form TE/TM modes, isotropic/uniaxial eps, isotropic mu
// -- mod_ModeSolver.jl:243-265
__device__ __forceinline__ void eigen_closed(const LayerDev &L, double w, double kx, double ky, cd *Q, cd *V) {
  const cd et = L.eps[0], ez = L.eps[8], mr = L.mu[0];
  const cd wc = mk(w);
  const double kr = hypot(kx, ky);
  // cos and sin of phi = atan(ky, kx), taken as kx / kr and ky / kr
  const double ikr = kr > 1e-300 ? rcp_nr(kr) : 1.0 / kr;
  const double cp = kr == 0.0 ? 1.0 : kx * ikr, sp = kr == 0.0 ? 0.0 : ky * ikr;
  const cd kr2 = mk(kr * kr);
  const cd kzo = csqrt(wc * wc * et * mr - kr2);
  // isotropic eps: kze equals kzo, taken as the same number
  const cd kze = L.ek == 0 ? kzo : csqrt(et * (wc * wc * mr - kr2 / ez));
  const cd A = mk(1.0) / (wc * mr);
  const cd Bn = mk(1.0) / (wc * et);
  const cd c = mk(cp), s = mk(sp);
  V[0] = -s; V[1] = c; V[2] = -(A * kzo * c); V[3] = -(A * kzo * s);
  V[4] = -s; V[5] = c; V[6] = A * kzo * c; V[7] = A * kzo * s;
  V[8] = Bn * kze * c; V[9] = Bn * kze * s; V[10] = -s; V[11] = c;
  V[12] = -(Bn * kze * c); V[13] = -(Bn * kze * s); V[14] = -s; V[15] = c;
  Q[0] = kzo; Q[1] = -kzo; Q[2] = kze; Q[3] = -kze;
}

// _modes_layer -- mod_ModeSolver.jl:274-326.  Output columns/entries are in
// slot order (f1, f2, b1, b2).  Returns false if the eigen solve failed or
// more than two modes fall on one side.
// `canon` (may be NULL) receives whether the layer took the closed form and its
// slots hold (TEf, TMf, TEb, TMb) unpermuted.
template <bool GENERAL = true>
__device__ __forceinline__ bool modes_layer(const LayerDev &L, double w, double kx, double ky, cd *Vm, cd *Qm,
                                            bool *canon = nullptr) {
  const double tol = 1.4901161193847656e-08;
  cd Qs[4], Vs[16];
  bool ok = true;
  const bool closed = (L.ek == 0 || L.ek == 1) && L.mk == 0;
  if (closed) {
    eigen_closed(L, w, kx, ky, Qs, Vs);
  } else {
    cd B[16];
    berreman(L, w, kx, ky, B);
    if (!GENERAL || is_block_form(L)) ok = eigen_block(B, Qs, Vs);
    else ok = eigen4(B, Qs, Vs);
  }
  // class per raw mode: 0 forward, 1 backward, 2 degenerate
  int cls[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    // Pz = 1/4 Re(V . R4s V) = 1/2 Re(conj(V1) V4 - conj(V2) V3)
    const double pz = 0.5 * ((Vs[4 * j].r * Vs[4 * j + 3].r + Vs[4 * j].i * Vs[4 * j + 3].i) -
                             (Vs[4 * j + 1].r * Vs[4 * j + 2].r + Vs[4 * j + 1].i * Vs[4 * j + 2].i));
    const double iq = Qs[j].i;
    int c = 2;
    if (fabs(pz) > tol) c = pz > 0 ? 0 : 1;
    else if (fabs(iq) > tol) c = iq > tol ? 0 : 1;
    cls[j] = c;
  }
  int nf = 0, nb = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    nf += cls[j] == 0;
    nb += cls[j] == 1;
  }
  ok = ok && nf <= 2 && nb <= 2;
  int dst[4];
  int cf = 0, cb = 0, cdg = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int dj;
    if (cls[j] == 0) dj = cf++;
    else if (cls[j] == 1) dj = 2 + cb++;
    else {
      const int i = cdg++;
      dj = (nf + i < 2) ? nf + i : 2 + nb + (i - (2 - nf));
    }
    dst[j] = dj;
  }
  // Move raw mode j to slot dst[j].  When every lane of the wave has the same
  // order and it is one of the usual two (closed form: TEf TMf | TEb TMb; the
  // (Re, Im)-sorted eigen solves: the two forward modes last), the moves are
  // fixed; otherwise each slot selects among the four modes.
  const int code = dst[0] | dst[1] << 2 | dst[2] << 4 | dst[3] << 6;
  const int wcode = __builtin_amdgcn_readfirstlane(code);
  constexpr int kClosed = 0 | 2 << 2 | 1 << 4 | 3 << 6, kSorted = 2 | 3 << 2 | 0 << 4 | 1 << 6;
  if (canon) *canon = closed && code == kClosed;
  if (__all(code == wcode) && (wcode == kClosed || wcode == kSorted)) {
    if (wcode == kClosed) {
#pragma unroll
      for (int slot = 0; slot < 4; ++slot) {
        const int j = slot == 0 ? 0 : slot == 1 ? 2 : slot == 2 ? 1 : 3;
        Qm[slot] = Qs[j];
#pragma unroll
        for (int r = 0; r < 4; ++r) Vm[4 * slot + r] = Vs[4 * j + r];
      }
    } else {
#pragma unroll
      for (int slot = 0; slot < 4; ++slot) {
        const int j = slot == 0 ? 2 : slot == 1 ? 3 : slot == 2 ? 0 : 1;
        Qm[slot] = Qs[j];
#pragma unroll
        for (int r = 0; r < 4; ++r) Vm[4 * slot + r] = Vs[4 * j + r];
      }
    }
    return ok;
  }
#pragma unroll
  for (int slot = 0; slot < 4; ++slot) {
    cd q = mk(0.0);
    cd v0 = mk(0.0), v1 = mk(0.0), v2 = mk(0.0), v3 = mk(0.0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool t = dst[j] == slot;
      q = sel(t, Qs[j], q);
      v0 = sel(t, Vs[4 * j], v0);
      v1 = sel(t, Vs[4 * j + 1], v1);
      v2 = sel(t, Vs[4 * j + 2], v2);
      v3 = sel(t, Vs[4 * j + 3], v3);
    }
    Qm[slot] = q;
    Vm[4 * slot] = v0; Vm[4 * slot + 1] = v1; Vm[4 * slot + 2] = v2; Vm[4 * slot + 3] = v3;
  }
  return ok;
}

// _match_modes (eigpenalty 0) of one side (cols c0, c0+1): C = Wa' Vb with
// W = R4s V.  Returns the swap decision and applies it.
__device__ __forceinline__ bool match_side(const cd *Vprev, cd *V, cd *Q, int c0) {
  cd C[4];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const cd *a = Vprev + 4 * (c0 + i);
      const cd *b = V + 4 * (c0 + j);
      // Wa = (a4, -a3, -a2, a1)
      cd s = conj(a[3]) * b[0];
      s = s - conj(a[2]) * b[1];
      s = s - conj(a[1]) * b[2];
      s = fmac_cj(s, a[0], b[3]);
      C[i + 2 * j] = s;
    }
  const double sid = abs2(C[0]) + abs2(C[3]);
  const double ssw = abs2(C[2]) + abs2(C[1]);
  const bool sw = ssw > sid;
  cswap_col(sw, Q, V, c0, c0 + 1);
  return sw;
}
__device__ __forceinline__ void apply_swaps(bool swf, bool swb, cd *V, cd *Q) {
  cswap_col(swf, Q, V, 0, 1);
  cswap_col(swb, Q, V, 2, 3);
}

// 4x4 LU with partial pivoting, X = A \ B (B 4xN).  Returns false on an
// exactly zero pivot.
template <int N>
__device__ __forceinline__ bool lu_solve4(cd *A, cd *B) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double amax = cabs(A[k + 4 * k]);
    int p = k;
#pragma unroll
    for (int i = k + 1; i < 4; ++i) {
      const double ai = cabs(A[i + 4 * k]);
      if (ai > amax) { amax = ai; p = i; }
    }
    ok = ok && amax != 0.0;
    // Row exchange: none when no lane pivots, fixed moves when every lane of
    // the wave pivots on the same row, per-lane selects otherwise.
    const int wp = __builtin_amdgcn_readfirstlane(p);
    const bool uniform = __all(p == wp);
    if (!(uniform && wp == k)) {
#pragma unroll
      for (int i = k + 1; i < 4; ++i) {
        if (uniform) {
          if (wp == i) {
#pragma unroll
            for (int j = k; j < 4; ++j) {
              const cd a = A[k + 4 * j];
              A[k + 4 * j] = A[i + 4 * j];
              A[i + 4 * j] = a;
            }
#pragma unroll
            for (int j = 0; j < N; ++j) {
              const cd a = B[k + 4 * j];
              B[k + 4 * j] = B[i + 4 * j];
              B[i + 4 * j] = a;
            }
          }
        } else {
          const bool t = p == i;
#pragma unroll
          for (int j = k; j < 4; ++j) {
            const cd a = A[k + 4 * j], b = A[i + 4 * j];
            A[k + 4 * j] = sel(t, b, a);
            A[i + 4 * j] = sel(t, a, b);
          }
#pragma unroll
          for (int j = 0; j < N; ++j) {
            const cd a = B[k + 4 * j], b = B[i + 4 * j];
            B[k + 4 * j] = sel(t, b, a);
            B[i + 4 * j] = sel(t, a, b);
          }
        }
      }
    }
    const cd rp = cinv(A[k + 4 * k]);
    A[k + 4 * k] = rp;  // the back substitution reads the reciprocal pivot
#pragma unroll
    for (int i = k + 1; i < 4; ++i) {
      const cd l = A[i + 4 * k] * rp;
#pragma unroll
      for (int j = k + 1; j < 4; ++j) A[i + 4 * j] = A[i + 4 * j] - l * A[k + 4 * j];
#pragma unroll
      for (int j = 0; j < N; ++j) B[i + 4 * j] = B[i + 4 * j] - l * B[k + 4 * j];
    }
  }
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = 3; i >= 0; --i) {
      cd s = B[i + 4 * j];
#pragma unroll
      for (int k = i + 1; k < 4; ++k) s = s - A[i + 4 * k] * B[k + 4 * j];
      B[i + 4 * j] = s * A[i + 4 * i];
    }
  return ok;
}

// The pivoted-QR minimum-norm solve that replaces an exactly singular LU in
// _S_Interface_Matrix (the SingularException fallback,
// mod_ScatteringMatrix.jl:35-44: `qr(L, ColumnNorm()) \ R`, the minimum-norm
// least-squares solution of the rank-r truncated system).  In place, B <- X.
// Column pivots by the largest downdated norm; the rank r is the first k with
// |R_kk| <= 4 eps |R_11|.  The minimum-norm solution of the r x 4 system
// R y = c comes from a Householder QR of R^H = Q2 [T; 0]: y = Q2 [T^-H c; 0].
// Per-lane pivots and ranks are selects; every index is a compile-time constant.
template <int N>
__device__ __forceinline__ void qr_minnorm4(cd *A, cd *B) {
  constexpr double kTol = 4.0 * 2.220446049250313e-16;
  int jp[4] = {0, 1, 2, 3};
  double cn[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) cn[j] = abs2(A[4 * j]) + abs2(A[4 * j + 1]) + abs2(A[4 * j + 2]) + abs2(A[4 * j + 3]);
  int rank = 4;
  bool act = true;
  double rmax = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int p = k;
    double cm = cn[k];
#pragma unroll
    for (int j = k + 1; j < 4; ++j) {
      const bool t = cn[j] > cm;
      cm = t ? cn[j] : cm;
      p = t ? j : p;
    }
#pragma unroll
    for (int j = k + 1; j < 4; ++j) {
      const bool t = act && p == j;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const cd a = A[i + 4 * k], b = A[i + 4 * j];
        A[i + 4 * k] = sel(t, b, a);
        A[i + 4 * j] = sel(t, a, b);
      }
      const double c0 = cn[k], c1 = cn[j];
      cn[k] = t ? c1 : c0;
      cn[j] = t ? c0 : c1;
      const int i0 = jp[k], i1 = jp[j];
      jp[k] = t ? i1 : i0;
      jp[j] = t ? i0 : i1;
    }
    double nrm2 = 0.0;
#pragma unroll
    for (int i = k; i < 4; ++i) nrm2 += abs2(A[i + 4 * k]);
    const double nrm = sqrt(nrm2);
    if (k == 0) rmax = nrm;
    const bool cut = !(nrm > kTol * fmax(rmax, 1e-300));
    rank = (act && cut) ? k : rank;
    act = act && !cut;
    // reflector H = I - 2 v v^H / (v^H v) with H x = alpha e_k, alpha = -|x| x_k / |x_k|
    const cd a0 = A[k + 4 * k];
    const double aa = cabs(a0);
    const cd alpha = rmul(-nrm, aa == 0.0 ? mk(1.0) : rmul(1.0 / aa, a0));
    cd v[4];
    v[k] = a0 - alpha;
#pragma unroll
    for (int i = k + 1; i < 4; ++i) v[i] = A[i + 4 * k];
    double vn = 0.0;
#pragma unroll
    for (int i = k; i < 4; ++i) vn += abs2(v[i]);
    const double f = (act && vn > 0.0) ? 2.0 / vn : 0.0;
#pragma unroll
    for (int j = k; j < 4; ++j) {
      cd t = mk(0.0);
#pragma unroll
      for (int i = k; i < 4; ++i) t = fmac_cj(t, v[i], A[i + 4 * j]);
      t = rmul(f, t);
#pragma unroll
      for (int i = k; i < 4; ++i) A[i + 4 * j] = A[i + 4 * j] - v[i] * t;
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cd t = mk(0.0);
#pragma unroll
      for (int i = k; i < 4; ++i) t = fmac_cj(t, v[i], B[i + 4 * j]);
      t = rmul(f, t);
#pragma unroll
      for (int i = k; i < 4; ++i) B[i + 4 * j] = B[i + 4 * j] - v[i] * t;
    }
#pragma unroll
    for (int j = k + 1; j < 4; ++j) cn[j] -= abs2(A[k + 4 * j]);
  }
  // H = R^H on the leading rank rows; its QR holds the reflectors in its lower
  // part (column k from row k), T's strict upper part above, T's diagonal in td
  cd H[16], td[4];
  double fk[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) H[j + 4 * i] = (i < rank && j >= i) ? conj(A[i + 4 * j]) : mk(0.0);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double nrm2 = 0.0;
#pragma unroll
    for (int i = k; i < 4; ++i) nrm2 += abs2(H[i + 4 * k]);
    const double nrm = sqrt(nrm2);
    const cd a0 = H[k + 4 * k];
    const double aa = cabs(a0);
    const cd alpha = rmul(-nrm, aa == 0.0 ? mk(1.0) : rmul(1.0 / aa, a0));
    cd v[4];
    v[k] = a0 - alpha;
#pragma unroll
    for (int i = k + 1; i < 4; ++i) v[i] = H[i + 4 * k];
    double vn = 0.0;
#pragma unroll
    for (int i = k; i < 4; ++i) vn += abs2(v[i]);
    const bool on = k < rank;
    const double f = (on && vn > 0.0) ? 2.0 / vn : 0.0;
#pragma unroll
    for (int j = k + 1; j < 4; ++j) {
      cd t = mk(0.0);
#pragma unroll
      for (int i = k; i < 4; ++i) t = fmac_cj(t, v[i], H[i + 4 * j]);
      t = rmul(f, t);
#pragma unroll
      for (int i = k; i < 4; ++i) H[i + 4 * j] = H[i + 4 * j] - v[i] * t;
    }
#pragma unroll
    for (int i = k; i < 4; ++i) H[i + 4 * k] = v[i];
    td[k] = on ? alpha : mk(1.0);
    fk[k] = f;
  }
#pragma unroll
  for (int j = 0; j < N; ++j) {
    cd y[4];
    // T^H w = c (T^H lower triangular: (T^H)_im = conj(T_mi) = conj(H_mi), m < i)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cd t = B[i + 4 * j];
#pragma unroll
      for (int m = 0; m < i; ++m) t = t - conj(H[m + 4 * i]) * y[m];
      y[i] = i < rank ? t * cinv(conj(td[i])) : mk(0.0);
    }
    // y = Q2 w, Q2 = H_0 H_1 ... H_{rank-1}
#pragma unroll
    for (int k = 3; k >= 0; --k) {
      cd t = mk(0.0);
#pragma unroll
      for (int i = k; i < 4; ++i) t = fmac_cj(t, H[i + 4 * k], y[i]);
      t = rmul(fk[k], t);
#pragma unroll
      for (int i = k; i < 4; ++i) y[i] = y[i] - H[i + 4 * k] * t;
    }
    // undo the column pivots: X[jp[i]] = y[i]
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      cd x = mk(0.0);
#pragma unroll
      for (int i = 0; i < 4; ++i) x = sel(jp[i] == r, y[i], x);
      B[r + 4 * j] = x;
    }
  }
}

// The closed form's mode coefficients: with u = (c, s), v = (-s, c) the TE
// columns are E = v, H = -+a u and the TM columns E = +-b u, H = v, where
// a = kzo / (w mu) and b = kze / (w eps_t) (mod_ModeSolver.jl:243-265).
__device__ __forceinline__ void closed_ab(const LayerDev &L, double w, const cd *Q, cd &a, cd &b) {
  a = Q[0] / (mk(w) * L.mu[0]);
  b = Q[1] / (mk(w) * L.eps[0]);
}

// _S_Interface_Matrix between two closed-form layers whose slots hold
// (TEf, TMf, TEb, TMb): TE and TM decouple, and with x = a (TE) or b (TM) the
// solve [-Vf_B Vb_A] [fB; bA] = [-Vf_A Vb_B] [fA; bB] gives, per polarisation,
// fB = (2 xA fA + (xB - xA) bB) / (xA + xB), bA = ((xA - xB) fA + 2 xB bB) / (xA + xB),
// so every 2x2 block of S is diagonal.
__device__ __forceinline__ void s_interface_closed(cd aA, cd bA, cd aB, cd bB, cd *S) {
  const cd ra = cinv(aA + aB), rb = cinv(bA + bB);
  const cd z = mk(0.0);
#pragma unroll
  for (int i = 0; i < 16; ++i) S[i] = z;
  S[0] = rmul(2.0, aA) * ra;       S[5] = rmul(2.0, bA) * rb;        // S11: fB <- fA
  S[8] = (aB - aA) * ra;           S[13] = (bB - bA) * rb;           // S12: fB <- bB
  S[2] = (aA - aB) * ra;           S[7] = (bA - bB) * rb;            // S21: bA <- fA
  S[10] = rmul(2.0, aB) * ra;      S[15] = rmul(2.0, bB) * rb;       // S22: bA <- bB
}

// The system of _S_Interface_Matrix: L = [-Vf_B  Vb_A], R = [-Vf_A  Vb_B].
__device__ __forceinline__ void interface_system(const cd *VA, const cd *VB, cd *L, cd *R) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    L[i] = -VB[i];
    L[8 + i] = VA[8 + i];
    R[i] = -VA[i];
    R[8 + i] = VB[8 + i];
  }
}

// _S_Interface_Matrix -- mod_ScatteringMatrix.jl:27-47:
// S = [-Vf_B  Vb_A] \ [-Vf_A  Vb_B] by LU.  Returns false on an exactly zero
// pivot (the reference's SingularException).
__device__ __forceinline__ bool s_interface(const cd *VA, const cd *VB, cd *S) {
  cd L[16];
  interface_system(VA, VB, L, S);
  return lu_solve4<4>(L, S);
}

// The same with the reference's answer to a SingularException (:35-44): lanes
// whose LU meets an exactly zero pivot rebuild the system and take the
// pivoted-QR minimum-norm solve.
__device__ __forceinline__ void s_interface_qr(const cd *VA, const cd *VB, cd *S) {
  if (!s_interface(VA, VB, S)) {
    cd L[16];
    interface_system(VA, VB, L, S);
    qr_minnorm4<4>(L, S);
  }
}

// The QR fallback of an interface solve in the passes that keep the previous
// layer's modes in LDS: one side's matched modes are re-solved (layer Lr with
// its recorded swaps), the other side's come from the lane's LDS slots
// (Vlds[i * 64], i = 0..15); lds_is_A says which side LDS holds.  S receives
// L \ R by the pivoted-QR minimum-norm solve.  Kept out of line so that its
// mode solve and QR do not add to the register pressure of the pass.
template <bool GENERAL>
__device__ __noinline__ void interface_fallback(const LayerDev *Lr, double w, double kx, double ky, bool swf, bool swb,
                                                const cd *Vlds, bool lds_is_A, cd *S) {
  cd Vr[16], Qr[4], Vl[16], L[16];
  modes_layer<GENERAL>(*Lr, w, kx, ky, Vr, Qr);
  apply_swaps(swf, swb, Vr, Qr);
#pragma unroll
  for (int i = 0; i < 16; ++i) Vl[i] = Vlds[i * 64];
  if (lds_is_A) interface_system(Vl, Vr, L, S);
  else interface_system(Vr, Vl, L, S);
  qr_minnorm4<4>(L, S);
}

// Redheffer star product -- mod_ScatteringMatrix.jl:65-85
__device__ __forceinline__ void redheffer(const cd *S1, const cd *S2, cd *S) {
  cd A11[4], A12[4], A21[4], A22[4], B11[4], B12[4], B21[4], B22[4];
  blk(S1, 0, 0, A11); blk(S1, 0, 1, A12); blk(S1, 1, 0, A21); blk(S1, 1, 1, A22);
  blk(S2, 0, 0, B11); blk(S2, 0, 1, B12); blk(S2, 1, 0, B21); blk(S2, 1, 1, B22);
  cd M1[4], M2[4], t[4], u[4], X11[4], X12[4], X21[4], X22[4];
  mm<2, 2, 2>(A12, B21, t);
  M1[0] = mk(1.0) - t[0]; M1[1] = -t[1]; M1[2] = -t[2]; M1[3] = mk(1.0) - t[3];
  mm<2, 2, 2>(B21, A12, t);
  M2[0] = mk(1.0) - t[0]; M2[1] = -t[1]; M2[2] = -t[2]; M2[3] = mk(1.0) - t[3];
  solve2<2>(M1, A11, t);
  mm<2, 2, 2>(B11, t, X11);
  solve2<2>(M1, A12, t);
  mm<2, 2, 2>(t, B22, u);
  mm<2, 2, 2>(B11, u, t);
#pragma unroll
  for (int i = 0; i < 4; ++i) X12[i] = B12[i] + t[i];
  solve2<2>(M2, B21, t);
  mm<2, 2, 2>(t, A11, u);
  mm<2, 2, 2>(A22, u, t);
#pragma unroll
  for (int i = 0; i < 4; ++i) X21[i] = A21[i] + t[i];
  solve2<2>(M2, B22, t);
  mm<2, 2, 2>(A22, t, X22);
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      S[i + 4 * j] = X11[i + 2 * j];
      S[i + 4 * (j + 2)] = X12[i + 2 * j];
      S[i + 2 + 4 * j] = X21[i + 2 * j];
      S[i + 2 + 4 * (j + 2)] = X22[i + 2 * j];
    }
}
// S1 * diag(p)  (S2 a propagation matrix)
__device__ __forceinline__ void redheffer_rdiag(const cd *S1, const cd *p, cd *S) {
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      S[i + 4 * j] = p[i] * S1[i + 4 * j];                                // B11 A11
      S[i + 4 * (j + 2)] = p[i] * (S1[i + 4 * (j + 2)] * p[2 + j]);        // B11 (A12 B22)
      S[i + 2 + 4 * j] = S1[i + 2 + 4 * j];                               // A21
      S[i + 2 + 4 * (j + 2)] = S1[i + 2 + 4 * (j + 2)] * p[2 + j];         // A22 B22
    }
}
// diag(p) * S2  (S1 a propagation matrix)
__device__ __forceinline__ void redheffer_ldiag(const cd *p, const cd *S2, cd *S) {
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      S[i + 4 * j] = S2[i + 4 * j] * p[j];                                // B11 A11
      S[i + 4 * (j + 2)] = S2[i + 4 * (j + 2)];                           // B12
      S[i + 2 + 4 * j] = p[2 + i] * (S2[i + 2 + 4 * j] * p[j]);            // A22 (B21 A11)
      S[i + 2 + 4 * (j + 2)] = p[2 + i] * S2[i + 2 + 4 * (j + 2)];         // A22 B22
    }
}

// _cond2x2_eq > 1/tol  (mod_FieldSolver.jl:33-53), tol = 1e-8
__device__ __forceinline__ bool ill_conditioned(const cd *A) {
  const double n1 = sqrt(abs2(A[0]) + abs2(A[1]));
  const double n2 = sqrt(abs2(A[2]) + abs2(A[3]));
  if (!(n1 > 0 && n2 > 0 && isfinite(n1) && isfinite(n2))) return true;
  const cd a0 = rmul(1.0 / n1, A[0]), a1 = rmul(1.0 / n1, A[1]);
  const cd a2 = rmul(1.0 / n2, A[2]), a3 = rmul(1.0 / n2, A[3]);
  // G = A' A
  const cd g00 = conj(a0) * a0 + conj(a1) * a1, g11 = conj(a2) * a2 + conj(a3) * a3;
  const cd g01 = conj(a0) * a2 + conj(a1) * a3, g10 = conj(a2) * a0 + conj(a3) * a1;
  const double tr = (g00 + g11).r;
  const double det = (g00 * g11 - g01 * g10).r;
  const double dd = tr * tr - 4.0 * det;
  const double disc = sqrt(dd > 0.0 ? dd : 0.0);
  const double l1 = (tr + disc) / 2.0, l2 = (tr - disc) / 2.0;
  if (l2 <= 0.0) return true;
  return sqrt(l1 / l2) > 1e8;
}

// Source_projection_Berreman_xyz for a closed-form source layer whose slots
// hold (TEf, TMf, TEb, TMb): (W'V) \ (W'J) = V^-1 J, and V^-1 splits by
// polarisation.  With J's components E_v, E_u, H_u, H_v along v = (-s, c),
// u = (c, s): f_TE = (E_v - H_u / a) / 2, b_TE = (E_v + H_u / a) / 2,
// f_TM = (H_v + E_u / b) / 2, b_TM = (H_v - E_u / b) / 2 (the closed form has
// eps_tz = 0, so J_Ht = R2s Jt).
__device__ __forceinline__ bool source_projection_closed(const LayerDev &L, double w, double kx, double ky,
                                                         const cd *Q, cd *sif, cd *sib) {
  const double kr = hypot(kx, ky);
  const double ikr = kr > 1e-300 ? rcp_nr(kr) : 1.0 / kr;
  const double c = kr == 0.0 ? 1.0 : kx * ikr, sn = kr == 0.0 ? 0.0 : ky * ikr;
  cd a, b;
  closed_ab(L, w, Q, a, b);
  const cd ia = cinv(a), ib = cinv(b);
  const cd mio = mk(0.0, -w);
  const cd jz = mio / (L.eps[8] * mk(w));  // J_Et of the z dipole is (kx, ky) jz
  // per moment q: (E_x, E_y, H_x, H_y) of J_Berreman = (0, 0, 0, -mio), (0, 0, mio, 0), (kx jz, ky jz, 0, 0)
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    cd Ev, Eu, Hu, Hv;
    if (q == 2) {
      Ev = mk(0.0); Eu = rmul(kr, jz); Hu = mk(0.0); Hv = mk(0.0);
    } else {
      const double hx = q == 0 ? 0.0 : 1.0, hy = q == 0 ? -1.0 : 0.0;
      Ev = mk(0.0); Eu = mk(0.0);
      Hu = rmul(c * hx + sn * hy, mio);
      Hv = rmul(-sn * hx + c * hy, mio);
    }
    const cd hu = Hu * ia, eu = Eu * ib;
    sif[2 * q] = rmul(0.5, Ev - hu);
    sib[2 * q] = rmul(0.5, Ev + hu);
    sif[2 * q + 1] = rmul(0.5, Hv + eu);
    sib[2 * q + 1] = rmul(0.5, Hv - eu);
  }
  // a = 0 or b = 0 (k_z = 0 on a light line) is the singular W'V of the reference
  return !(a.r == 0.0 && a.i == 0.0) && !(b.r == 0.0 && b.i == 0.0);
}

// Source_projection_Berreman_xyz -- mod_SourceProj.jl:34-60.  sif, sib:
// 2x3 column-major (moments x, y, z).
__device__ __forceinline__ bool source_projection(const LayerDev &L, double w, double kx, double ky, const cd *V,
                                                  cd *sif, cd *sib) {
  const cd etz1 = L.eps[6], etz2 = L.eps[7], ezz = L.eps[8];
  const cd mio = mk(0.0, -w);
  const cd den = ezz * mk(w);
  cd J[12];
  // column q = 0, 1: tangential dipoles; q = 2: z dipole
  const cd zr = rmul(0.0, mio);
  J[0] = mk(kx) * zr / den; J[1] = mk(ky) * zr / den; J[2] = zr - etz2 * zr / ezz; J[3] = -mio + etz1 * zr / ezz;
  J[4] = mk(kx) * zr / den; J[5] = mk(ky) * zr / den; J[6] = mio - etz2 * zr / ezz; J[7] = -zr + etz1 * zr / ezz;
  J[8] = mk(kx) * mio / den; J[9] = mk(ky) * mio / den; J[10] = zr - etz2 * mio / ezz; J[11] = -zr + etz1 * mio / ezz;
  // A = W' V, B = W' J with W = R4s V: W_col = (v4, -v3, -v2, v1)
  cd A[16], B[12];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const cd *a = V + 4 * i;
    const cd w0 = a[3], w1 = -a[2], w2 = -a[1], w3 = a[0];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const cd *b = V + 4 * j;
      cd s = conj(w0) * b[0];
      s = fmac_cj(s, w1, b[1]);
      s = fmac_cj(s, w2, b[2]);
      s = fmac_cj(s, w3, b[3]);
      A[i + 4 * j] = s;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const cd *b = J + 4 * j;
      cd s = conj(w0) * b[0];
      s = fmac_cj(s, w1, b[1]);
      s = fmac_cj(s, w2, b[2]);
      s = fmac_cj(s, w3, b[3]);
      B[i + 4 * j] = s;
    }
  }
  const bool ok = lu_solve4<3>(A, B);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    sif[2 * q] = B[4 * q]; sif[2 * q + 1] = B[4 * q + 1];
    sib[2 * q] = B[4 * q + 2]; sib[2 * q + 1] = B[4 * q + 3];
  }
  return ok;
}

// mode-field basis of one layer (_mode_field_core, _fill_obs_basis!,
// mod_CavityTransfer.jl:87-130): out[0..11] = Exf Eyf Ezf Exb Eyb Ezb (2 modes
// each), out[12..15] = KZf, KZb.  A mode with a non-finite value stays zero.
__device__ __forceinline__ void obs_basis(const LayerDev &L, double w, double kx, double ky, const cd *V, const cd *Q,
                                          cd *out) {
  const cd ezt1 = L.eps[2], ezt2 = L.eps[5], ezz = L.eps[8];
  const cd iw = mk(1.0) / mk(w);
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    cd vals[8];
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
      const cd *v = V + 4 * (2 * dir + m);
      const cd ez = -((ezt1 * v[0] + ezt2 * v[1] + iw * (mk(kx) * v[3] - mk(ky) * v[2])) / ezz);
      vals[3 * dir + 0] = v[0];
      vals[3 * dir + 1] = v[1];
      vals[3 * dir + 2] = ez;
      vals[6 + dir] = Q[2 * dir + m];
    }
    bool fin = true;
#pragma unroll
    for (int i = 0; i < 8; ++i) fin = fin && isfin(vals[i]);
    const cd z = mk(0.0);
    out[0 + m] = sel(fin, vals[0], z);
    out[2 + m] = sel(fin, vals[1], z);
    out[4 + m] = sel(fin, vals[2], z);
    out[6 + m] = sel(fin, vals[3], z);
    out[8 + m] = sel(fin, vals[4], z);
    out[10 + m] = sel(fin, vals[5], z);
    out[12 + m] = sel(fin, vals[6], z);
    out[14 + m] = sel(fin, vals[7], z);
  }
}

// clamped exp(i kz z) -- mod_ClosedFormField.jl:59-69
__device__ __forceinline__ cd phase(cd kz, double z) {
  const double a = -kz.i * z, b = kz.r * z;
  if (!isfinite(a) || !isfinite(b) || a > 700.0 || a < -700.0) return mk(0.0);
  return cexp(cd{a, b});
}
__device__ __forceinline__ cd fin0(cd z) { return isfin(z) ? z : mk(0.0); }

// Box sources (mod_ClosedFormField.jl:86-91, 790-796).
// _boxphase: [e^{iQ(l+dz/2)} - e^{iQ(l-dz/2)}] / (iQ dz), 4th-order Taylor
// guard below |Q dz| = 1e-3.
__device__ __noinline__ cd boxphase(cd Q, double l, double dz) {
  const cd y = rmul(dz, Q);
  if (cabs(y) < 1e-3) {
    const cd y2 = y * y;
    return phase(Q, l) * (mk(1.0) - rmul(1.0 / 24.0, y2) + rmul(1.0 / 1920.0, y2 * y2));
  }
  return (phase(Q, l + 0.5 * dz) - phase(Q, l - 0.5 * dz)) / times_i(y);
}
// _ramp: expm1(iQa) / (iQ dz), with expm1(x + it) = expm1(x) - 2 e^x sin^2(t/2)
// + i e^x sin t (Julia's complex expm1)
__device__ __noinline__ cd ramp(cd Q, double a, double dz) {
  if (cabs(Q) * a < 1e-12) return mk(a / dz);
  const cd y = times_i(rmul(a, Q));
  const cd den = times_i(rmul(dz, Q));
  if (y.r < -700.0) return mk(-1.0) / den;
  const double em1 = expm1(y.r);
  cd e;
  if (y.i == 0.0) {
    e = cd{em1, 0.0};
  } else {
    const double er = em1 + 1.0, sh = sin(0.5 * y.i);
    e = cd{em1 - 2.0 * er * sh * sh, er * sin(y.i)};
  }
  return e / den;
}

// ---------------------------------------------------------------------------
// The node kernel
// ---------------------------------------------------------------------------
// CASE_TOP: source and observation both in the first layer (the source above
// the stack).  Then RL = 0 and only the reflection block S21 of SRight[0] is
// needed; it follows the bottom-up recursion
//   R_k = eff21 + eff22 ((I - R_{k+1} eff12) \ R_{k+1}) eff11,  R_{N-1} = 0,
// the S21 block of SRight[k] = eff_k * SRight[k+1] (mod_ScatteringMatrix.jl:65-85,
// 130-135), with a 2x2 running state and no mode matching (layer 1's order is
// never changed by _match_modes, and the deeper layers' order cancels in R).
enum { CASE_SAME = 0, CASE_RIGHT = 1, CASE_LEFT = 2, CASE_TOP = 3 };

// Per-lane LDS area of the general cases, [slot][64 lanes] (conflict-free):
// the previous layer's mode matrix during the passes, and the 2x2 cascade
// blocks that the cavity and the walk read afterwards.  Keeping them out of
// registers lets a lane carry a whole layer solve without spilling.
struct LaneMat {
  cd *p;
  int lane;
  __device__ __forceinline__ cd get(int i) const { return p[i * 64 + lane]; }
  __device__ __forceinline__ void set(int i, cd v) const { p[i * 64 + lane] = v; }
  __device__ __forceinline__ void put4(int i, const cd *b) const {
#pragma unroll
    for (int k = 0; k < 4; ++k) set(i + k, b[k]);
  }
  __device__ __forceinline__ void get4(int i, cd *b) const {
#pragma unroll
    for (int k = 0; k < 4; ++k) b[k] = get(i + k);
  }
};
// slots: V 0-15; RL = SLeft[s]_12; SLs22 = SLeft[s]_22; Rs = SRight[s]_21;
// Ts = SRight[s]_11; X0..X2: CASE_RIGHT SRight[o]_11, SRight[o+1]_11,
// SRight[o+1]_21 -- CASE_LEFT SLeft[o]_12, SLeft[o]_22, SLeft[o+1]_22;
// Sint[o] blocks 21 and 22.
enum { SL_V = 0, SL_RL = 16, SL_SLS22 = 20, SL_RS = 24, SL_TS = 28, SL_X0 = 32, SL_X1 = 36, SL_X2 = 40,
       SL_SI21 = 44, SL_SI22 = 48, SL_END = 52 };
template <int CASE>
struct LdsSlots {
  static constexpr int n = CASE == CASE_TOP ? 1 : (CASE == CASE_SAME ? 32 : SL_END);
};

// _match_modes (eigpenalty 0) of one side against the previous layer in LDS.
__device__ __forceinline__ bool match_side_lds(const LaneMat &Vp, cd *V, cd *Q, int c0) {
  cd C[4];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int a = 4 * (c0 + i);
      const cd *b = V + 4 * (c0 + j);
      cd t = conj(Vp.get(a + 3)) * b[0];
      t = t - conj(Vp.get(a + 2)) * b[1];
      t = t - conj(Vp.get(a + 1)) * b[2];
      t = fmac_cj(t, Vp.get(a), b[3]);
      C[i + 2 * j] = t;
    }
  const bool sw = abs2(C[2]) + abs2(C[1]) > abs2(C[0]) + abs2(C[3]);
  cswap_col(sw, Q, V, c0, c0 + 1);
  return sw;
}

__device__ __forceinline__ void prop4(const cd *Q, double d, cd *p) {
  p[0] = expi(Q[0], d);
  p[1] = expi(Q[1], d);
  p[2] = expi(-Q[2], d);
  p[3] = expi(-Q[3], d);
}

// The cascade of the general cases in two lean passes (same algebra as
// S_all_Matrix, mod_ScatteringMatrix.jl:112-154, restricted to the blocks the
// cavity and the walk use):
//   forward, j = 0..N-1: modes with continuity matching (the swap decisions
//     are recorded per layer); the SLeft prefix needs only the blocks 12 and 22
//     of A = SLeft[j] * Sprop[j], since (A * B)_12 = B12 + B11 (I - A12 B21)^-1
//     A12 B22 and (A * B)_22 = A22 (I - B21 A12)^-1 B22;
//   backward, k = N-2..s: SRight[k] = eff_k * SRight[k+1] with
//     eff_k = Sprop[k] * Sint[k]; its blocks 21 (R) and 11 (T) follow the 2x2
//     recursions R_k = e21 + e22 ((I - R e12) \ R) e11 and
//     T_k = T_{k+1} ((I - e12 R) \ e11) from R = 0, T = I at k = N-1, with
//     the modes re-solved in the recorded matched order.
// Results go to the lane's LDS slots; returns the recorded swap bits
// (bit 2j: layer j's forward pair swapped, bit 2j+1: its backward pair).
template <int CASE, bool GENERAL, bool ROBUST>
__device__ uint64_t cascade_general(const NodeParams &P, double kx, double ky, const LaneMat &st, bool &ok,
                                    bool &redo) {
  const double w = P.omega;
  const int N = P.nl, s = P.src, o = P.obs;
  uint64_t bits = 0;
  {
    cd L12[4], L22[4];
    for (int j = 0; j < N; ++j) {
      const LayerDev &L = P.L[j];
      cd Q[4], Lm[16], Si[16];
      const bool need_int = j > 0 && (j <= s || (CASE != CASE_SAME && j == o + 1));
      {
        cd V[16];
        ok = modes_layer<GENERAL>(L, w, kx, ky, V, Q) && ok;
        if (j > 0) {
          const bool sf = match_side_lds(st, V, Q, 0);
          const bool sb = match_side_lds(st, V, Q, 2);
          bits |= ((uint64_t)sf << (2 * j)) | ((uint64_t)sb << (2 * j + 1));
          if (need_int) {
            // _S_Interface_Matrix: [-Vf_B  Vb_A] \ [-Vf_A  Vb_B], A = layer j-1
#pragma unroll
            for (int i = 0; i < 8; ++i) {
              Lm[i] = -V[i];
              Lm[8 + i] = st.get(SL_V + 8 + i);
              Si[i] = -st.get(SL_V + i);
              Si[8 + i] = V[8 + i];
            }
          }
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) st.set(SL_V + i, V[i]);
      }
      if (j == 0) {
        // SLeft[0] = I;  A = I * Sprop[0]
        if (s == 0) {
          const cd z[4] = {mk(0.0), mk(0.0), mk(0.0), mk(0.0)}, id[4] = {mk(1.0), mk(0.0), mk(0.0), mk(1.0)};
          st.put4(SL_RL, z);
          st.put4(SL_SLS22, id);
        }
        cd p[4];
        prop4(Q, L.d, p);
#pragma unroll
        for (int i = 0; i < 4; ++i) L12[i] = mk(0.0);
        L22[0] = p[2]; L22[1] = mk(0.0); L22[2] = mk(0.0); L22[3] = p[3];
      } else if (need_int) {
        const bool lu_ok = lu_solve4<4>(Lm, Si);  // Si = Sint[j-1]
        if (!ROBUST) redo = redo || !lu_ok;
        else if (!lu_ok) {
          // SingularException: layer j-1 re-solved in its matched order,
          // layer j from LDS, and the QR fallback
          cd Sf[16];
          interface_fallback<GENERAL>(&P.L[j - 1], w, kx, ky, (bits >> (2 * (j - 1))) & 1,
                                      (bits >> (2 * (j - 1) + 1)) & 1, st.p + SL_V * 64 + st.lane, false, Sf);
#pragma unroll
          for (int i = 0; i < 16; ++i) Si[i] = Sf[i];
        }
        if (j <= s) {
          cd B11[4], B12[4], B21[4], B22[4];
          blk(Si, 0, 0, B11); blk(Si, 0, 1, B12); blk(Si, 1, 0, B21); blk(Si, 1, 1, B22);
          cd t[4], M[4], X[4], Y[4], S12[4], S22[4];
          mm<2, 2, 2>(L12, B21, t);
          M[0] = mk(1.0) - t[0]; M[1] = -t[1]; M[2] = -t[2]; M[3] = mk(1.0) - t[3];
          solve2<2>(M, L12, X);
          mm<2, 2, 2>(X, B22, Y);
          mm<2, 2, 2>(B11, Y, t);
#pragma unroll
          for (int i = 0; i < 4; ++i) S12[i] = B12[i] + t[i];
          mm<2, 2, 2>(B21, L12, t);
          M[0] = mk(1.0) - t[0]; M[1] = -t[1]; M[2] = -t[2]; M[3] = mk(1.0) - t[3];
          solve2<2>(M, B22, X);
          mm<2, 2, 2>(L22, X, S22);
          if (j == s) {
            st.put4(SL_RL, S12);
            st.put4(SL_SLS22, S22);
          }
          if (CASE == CASE_LEFT) {
            if (j == o) { st.put4(SL_X0, S12); st.put4(SL_X1, S22); }
            if (j == o + 1) st.put4(SL_X2, S22);
          }
          if (j < s) {
            cd p[4];
            prop4(Q, L.d, p);
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
              for (int r = 0; r < 2; ++r) {
                L12[r + 2 * c] = p[r] * (S12[r + 2 * c] * p[2 + c]);
                L22[r + 2 * c] = S22[r + 2 * c] * p[2 + c];
              }
          }
        }
        if (CASE != CASE_SAME && j == o + 1) {
          cd b[4];
          blk(Si, 1, 0, b);
          st.put4(SL_SI21, b);
          blk(Si, 1, 1, b);
          st.put4(SL_SI22, b);
        }
      }
    }
  }
  {
    cd R[4] = {mk(0.0), mk(0.0), mk(0.0), mk(0.0)}, T[4] = {mk(1.0), mk(0.0), mk(0.0), mk(1.0)};
    if (CASE == CASE_RIGHT) {
      if (o == N - 1) st.put4(SL_X0, T);
      if (o + 1 == N - 1) { st.put4(SL_X1, T); st.put4(SL_X2, R); }
    }
    {
      cd V[16], Q[4];
      modes_layer<GENERAL>(P.L[N - 1], w, kx, ky, V, Q);
      apply_swaps((bits >> (2 * (N - 1))) & 1, (bits >> (2 * (N - 1) + 1)) & 1, V, Q);
#pragma unroll
      for (int i = 0; i < 16; ++i) st.set(SL_V + i, V[i]);
    }
    for (int kk = N - 2; kk >= s; --kk) {
      const LayerDev &L = P.L[kk];
      cd Q[4], Lm[16], Si[16];
      {
        cd V[16];
        modes_layer<GENERAL>(L, w, kx, ky, V, Q);
        apply_swaps((bits >> (2 * kk)) & 1, (bits >> (2 * kk + 1)) & 1, V, Q);
        // Sint[kk]: A = layer kk, B = layer kk+1 (in LDS)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          Lm[i] = -st.get(SL_V + i);
          Lm[8 + i] = V[8 + i];
          Si[i] = -V[i];
          Si[8 + i] = st.get(SL_V + 8 + i);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) st.set(SL_V + i, V[i]);
      }
      const bool lu_ok = lu_solve4<4>(Lm, Si);
      if (!ROBUST) redo = redo || !lu_ok;
      else if (!lu_ok) {
        // SingularException: layer kk from LDS, layer kk+1 re-solved in its
        // matched order, and the QR fallback
        cd Sf[16];
        interface_fallback<GENERAL>(&P.L[kk + 1], w, kx, ky, (bits >> (2 * (kk + 1))) & 1,
                                    (bits >> (2 * (kk + 1) + 1)) & 1, st.p + SL_V * 64 + st.lane, true, Sf);
#pragma unroll
        for (int i = 0; i < 16; ++i) Si[i] = Sf[i];
      }
      cd p[4];
      prop4(Q, L.d, p);
      cd e11[4], e12[4], e21[4], e22[4];
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          e11[r + 2 * c] = Si[r + 4 * c] * p[c];
          e12[r + 2 * c] = Si[r + 4 * (c + 2)];
          e21[r + 2 * c] = p[2 + r] * (Si[r + 2 + 4 * c] * p[c]);
          e22[r + 2 * c] = p[2 + r] * Si[r + 2 + 4 * (c + 2)];
        }
      cd t[4], M[4], X[4], Y[4], Z[4];
      mm<2, 2, 2>(R, e12, t);
      M[0] = mk(1.0) - t[0]; M[1] = -t[1]; M[2] = -t[2]; M[3] = mk(1.0) - t[3];
      solve2<2>(M, R, X);
      mm<2, 2, 2>(X, e11, Y);
      mm<2, 2, 2>(e22, Y, Z);
      mm<2, 2, 2>(e12, R, t);
#pragma unroll
      for (int i = 0; i < 4; ++i) R[i] = e21[i] + Z[i];
      M[0] = mk(1.0) - t[0]; M[1] = -t[1]; M[2] = -t[2]; M[3] = mk(1.0) - t[3];
      solve2<2>(M, e11, X);
      mm<2, 2, 2>(T, X, Y);
#pragma unroll
      for (int i = 0; i < 4; ++i) T[i] = Y[i];
      if (kk == s) { st.put4(SL_RS, R); st.put4(SL_TS, T); }
      if (CASE == CASE_RIGHT) {
        if (kk == o) st.put4(SL_X0, T);
        if (kk == o + 1) { st.put4(SL_X1, T); st.put4(SL_X2, R); }
      }
    }
  }
  return bits;
}

// Redo lists.  A fast solve that meets an exactly singular interface marks
// its node (one byte per node from byte 8; word 0 "some node is marked", word 1
// a block counter); a redo launch of kRedoBlocks blocks solves the marked
// nodes again with the reference's pivoted-QR fallback, clears their marks,
// and its last block clears the two words.  Plain stores (no atomics) keep the
// fast kernels' register allocation as it is.
constexpr int kRedoBlocks = 16;
__device__ __forceinline__ void redo_mark(int *list, int64_t k) {
  ((unsigned char *)(list + 2))[k] = 1;
  list[0] = 1;
}
__device__ __forceinline__ bool redo_any(const int *list) { return __atomic_load_n(list, __ATOMIC_RELAXED) != 0; }
__device__ __forceinline__ bool redo_take(int *list, int64_t k) {
  unsigned char *m = (unsigned char *)(list + 2) + k;
  if (!*m) return false;
  *m = 0;
  return true;
}
// The redo launch's last block clears the two words.
__device__ __forceinline__ void redo_finish(int *list) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(list + 1, 1) == (int)gridDim.x - 1) {
      list[0] = 0;
      list[1] = 0;
    }
  }
}

// One Green k-node.  ROBUST = false is the fast solve: an interface LU that
// meets an exactly zero pivot only sets `redo`, and the caller solves that
// node again with ROBUST = true, which takes the reference's pivoted-QR
// fallback there (mod_ScatteringMatrix.jl:35-44).  Keeping the fallback out of
// the fast solve leaves its register allocation as it is.
template <int CASE, bool GENERAL, bool ROBUST>
__device__ void solve_node(const NodeParams &P, int64_t k, const LaneMat &st, bool &redo, int pair0, int pair1) {
  if (P.rows) k = P.rows[k];
  const double kx = P.kx[k], ky = P.ky[k], w = P.omega;
  const int N = P.nl, s = P.src, o = P.obs;
  bool ok = true, sing = false;
  cd RL[4], SLs22[4];  // SLeft[s] blocks (0,1) and (1,1)
  cd Rs[4], Ts[4];     // SRight[s] blocks (1,0) and (0,0)
  bool swS_f = false, swS_b = false, swO_f = false, swO_b = false;
  if (CASE == CASE_TOP) {
#pragma unroll
    for (int i = 0; i < 4; ++i) { RL[i] = mk(0.0); SLs22[i] = mk(i == 0 || i == 3 ? 1.0 : 0.0); Ts[i] = mk(0.0); }
    cd R[4] = {mk(0.0), mk(0.0), mk(0.0), mk(0.0)};
    cd Vn[16], Qn[4];
    bool canon_n = false;
    ok = modes_layer<GENERAL>(P.L[N - 1], w, kx, ky, Vn, Qn, &canon_n) && ok;
    for (int j = N - 2; j >= 0; --j) {
      const LayerDev &L = P.L[j];
      cd V[16], Q[4];
      bool canon = false;
      ok = modes_layer<GENERAL>(L, w, kx, ky, V, Q, &canon) && ok;
      bool regular = true;
      if (ROBUST) {
        // x_A + x_B = 0 is the exactly singular closed-form interface: the
        // general solve below takes the reference's LU and its QR fallback
        cd aA, bA, aB, bB;
        closed_ab(L, w, Q, aA, bA);
        closed_ab(P.L[j + 1], w, Qn, aB, bB);
        const cd sa = aA + aB, sb = bA + bB;
        regular = (sa.r != 0.0 || sa.i != 0.0) && (sb.r != 0.0 || sb.i != 0.0);
      }
      if (__all(canon && canon_n && regular)) {
        // both layers in the closed form's slot order: TE and TM decouple and
        // every block of the interface S and of the layer's (S_int, S_prop)
        // product is diagonal (per polarisation c: TE 0, TM 1)
        cd aA, bA, aB, bB;
        closed_ab(L, w, Q, aA, bA);
        closed_ab(P.L[j + 1], w, Qn, aB, bB);
        cd Sint[16];
        s_interface_closed(aA, bA, aB, bB, Sint);
        const bool fin = isfin(Sint[0]) && isfin(Sint[5]) && isfin(Sint[8]) && isfin(Sint[13]);
        if (ROBUST) sing = !fin || sing;
        else redo = redo || !fin;
        cd p[2];
        p[0] = expi(Q[0], L.d);
        if (L.ek == 0) p[1] = p[0];
        else p[1] = expi(Q[1], L.d);
        const cd e11[2] = {Sint[0] * p[0], Sint[5] * p[1]};
        const cd e12[2] = {Sint[8], Sint[13]};
        const cd e21[2] = {p[0] * (Sint[2] * p[0]), p[1] * (Sint[7] * p[1])};
        const cd e22[2] = {p[0] * Sint[10], p[1] * Sint[15]};
        // R <- e21 + e22 ((I - R e12) \ R) e11
        cd M[4], X[4];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int r = 0; r < 2; ++r) M[r + 2 * c] = mk(r == c ? 1.0 : 0.0) - R[r + 2 * c] * e12[c];
        solve2<2>(M, R, X);
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int r = 0; r < 2; ++r) R[r + 2 * c] = e22[r] * (X[r + 2 * c] * e11[c]);
        R[0] = e21[0] + R[0];
        R[3] = e21[1] + R[3];
      } else {
        cd Sint[16];
        if (ROBUST) s_interface_qr(V, Vn, Sint);
        else redo = !s_interface(V, Vn, Sint) || redo;
        // the closed form's backward slots hold -Q of the forward ones, so their
        // phases e^{-i Qb d} are the forward phases (an isotropic layer has one)
        cd p[4];
        p[0] = expi(Q[0], L.d);
        if (__all(canon)) {
          if (L.ek == 0) p[1] = p[0];
          else p[1] = expi(Q[1], L.d);
          p[2] = p[0];
          p[3] = p[1];
        } else {
          p[1] = expi(Q[1], L.d);
          p[2] = expi(-Q[2], L.d);
          p[3] = expi(-Q[3], L.d);
        }
        cd e11[4], e12[4], e21[4], e22[4];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            e11[r + 2 * c] = Sint[r + 4 * c] * p[c];
            e12[r + 2 * c] = Sint[r + 4 * (c + 2)];
            e21[r + 2 * c] = p[2 + r] * (Sint[r + 2 + 4 * c] * p[c]);
            e22[r + 2 * c] = p[2 + r] * Sint[r + 2 + 4 * (c + 2)];
          }
        cd t[4], M[4], X[4], Y[4], Z[4];
        mm<2, 2, 2>(R, e12, t);
        M[0] = mk(1.0) - t[0]; M[1] = -t[1]; M[2] = -t[2]; M[3] = mk(1.0) - t[3];
        solve2<2>(M, R, X);
        mm<2, 2, 2>(X, e11, Y);
        mm<2, 2, 2>(e22, Y, Z);
#pragma unroll
        for (int i = 0; i < 4; ++i) R[i] = e21[i] + Z[i];
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) Vn[i] = V[i];
#pragma unroll
      for (int i = 0; i < 4; ++i) Qn[i] = Q[i];
      canon_n = canon;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) Rs[i] = R[i];
  } else {
    const uint64_t bits = cascade_general<CASE, GENERAL, ROBUST>(P, kx, ky, st, ok, redo);
    swS_f = (bits >> (2 * s)) & 1;
    swS_b = (bits >> (2 * s + 1)) & 1;
    swO_f = (bits >> (2 * o)) & 1;
    swO_b = (bits >> (2 * o + 1)) & 1;
    st.get4(SL_RL, RL);
    st.get4(SL_RS, Rs);
    st.get4(SL_TS, Ts);
    if (CASE == CASE_LEFT) st.get4(SL_SLS22, SLs22);
    else {
#pragma unroll
      for (int i = 0; i < 4; ++i) SLs22[i] = mk(0.0);
    }
  }

  // ---- source layer: modes (matched order), seeds ---------------------------
  const LayerDev &Ls = P.L[s];
  cd Vs[16], Qs[4];
  // The forward pass applied, at layer s, only the swap relative to layer s-1's
  // matched order; replaying the raw solve with the recorded decisions
  // reproduces the matched order of layer s exactly.
  bool canon_s = false;
  modes_layer<GENERAL>(Ls, w, kx, ky, Vs, Qs, &canon_s);
  apply_swaps(swS_f, swS_b, Vs, Qs);
  cd sif[6], sib[6];
  if (__all(canon_s && !swS_f && !swS_b)) sing = !source_projection_closed(Ls, w, kx, ky, Qs, sif, sib) || sing;
  else sing = !source_projection(Ls, w, kx, ky, Vs, sif, sib) || sing;
  bool seeds_fin = true;
#pragma unroll
  for (int i = 0; i < 6; ++i) seeds_fin = seeds_fin && isfin(sif[i]) && isfin(sib[i]);

  // ---- cavity of the source layer (_R_cavity, _cavity_resummation) ---------
  const double ds = Ls.d;
  // RRr = diag(ph) * SRight[s] (_R_cavity): its blocks 21 (RR) and 11
  cd RR[4], RRr11[4];
  // closed-form source layer in slot order: Qb = -Qf, so e^{i Qb d} = e^{-i Qf d}
  const bool phase_pairs = __all(canon_s && !swS_f && !swS_b);
  {
    cd ph[4];
    ph[0] = expi(Qs[0], -ds);
    if (phase_pairs) {
      if (Ls.ek == 0) ph[1] = ph[0];
      else ph[1] = expi(Qs[1], -ds);
      ph[2] = ph[0];
      ph[3] = ph[1];
    } else {
      ph[1] = expi(Qs[1], -ds);
      ph[2] = expi(Qs[2], ds);
      ph[3] = expi(Qs[3], ds);
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        RR[r + 2 * c] = ph[2 + r] * (Rs[r + 2 * c] * ph[c]);
        RRr11[r + 2 * c] = Ts[r + 2 * c] * ph[c];
      }
  }
  // F+ = (I - PF RL PB RR) \ (E+ + PF RL E-),  F- = E- + PB RR F+   (2x4 each)
  cd Fp[8], Fm[8];
  {
    cd PF[2], PB[2];
    PF[0] = expi(Qs[0], ds);
    if (phase_pairs) {
      if (Ls.ek == 0) PF[1] = PF[0];
      else PF[1] = expi(Qs[1], ds);
      PB[0] = PF[0];
      PB[1] = PF[1];
    } else {
      PF[1] = expi(Qs[1], ds);
      PB[0] = expi(Qs[2], -ds);
      PB[1] = expi(Qs[3], -ds);
    }
    cd t[4], A[4];
    // (PF*RL)*PB
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) t[i + 2 * j] = (PF[i] * RL[i + 2 * j]) * PB[j];
    cd u[4];
    mm<2, 2, 2>(t, RR, u);
    A[0] = mk(1.0) - u[0]; A[1] = -u[1]; A[2] = -u[2]; A[3] = mk(1.0) - u[3];
    cd rhs[8];
    // E+ = [I 0], E- = [0 I];  RL*E- = [0 RL]
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const cd ep = mk((j == i) ? 1.0 : 0.0);
        const cd rle = (j >= 2) ? RL[i + 2 * (j - 2)] : mk(0.0);
        rhs[i + 2 * j] = ep + PF[i] * rle;
      }
    solve2<4>(A, rhs, Fp);
    cd rf[8];
    mm<2, 2, 4>(RR, Fp, rf);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const cd em = mk((j - 2 == i) ? 1.0 : 0.0);
        Fm[i + 2 * j] = em + PB[i] * rf[i + 2 * j];
      }
  }

  // ---- T^o (_cavity_transfer_columns / _walk_stack) -------------------------
  cd T[16];
  cd Vo[16], Qo[4];
  if (CASE == CASE_SAME || CASE == CASE_TOP) {
    // [RL F- ; RR F+]
    cd a[8], b[8];
    mm<2, 2, 4>(RL, Fm, a);
    mm<2, 2, 4>(RR, Fp, b);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      T[4 * j] = a[2 * j]; T[4 * j + 1] = a[2 * j + 1];
      T[4 * j + 2] = b[2 * j]; T[4 * j + 3] = b[2 * j + 1];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) Vo[i] = Vs[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) Qo[i] = Qs[i];
  } else {
    const LayerDev &Lo = P.L[o];
    modes_layer<GENERAL>(Lo, w, kx, ky, Vo, Qo);
    apply_swaps(swO_f, swO_b, Vo, Qo);
    cd Ef[8], Ebnext[8];
    if (CASE == CASE_RIGHT) {
      cd cp[8];
      mm<2, 2, 4>(RRr11, Fp, cp);  // c- = 0
      if (o == N - 1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) { Ef[i] = flushd(cp[i]); Ebnext[i] = mk(0.0); }
      } else {
        cd S11[4];
        st.get4(SL_X0, S11);
        if (ill_conditioned(S11)) {
#pragma unroll
          for (int i = 0; i < 8; ++i) Ef[i] = mk(0.0);
        } else {
          cd bp[8];
          solve2<4>(S11, cp, bp);
#pragma unroll
          for (int i = 0; i < 8; ++i) Ef[i] = flushd(bp[i]);
        }
        if (o + 1 == N - 1) {
#pragma unroll
          for (int i = 0; i < 8; ++i) Ebnext[i] = mk(0.0);
        } else {
          cd S11b[4], S21b[4], bp[8], bm[8];
          st.get4(SL_X1, S11b);
          st.get4(SL_X2, S21b);
          if (ill_conditioned(S11b)) {
#pragma unroll
            for (int i = 0; i < 8; ++i) bp[i] = mk(0.0);
          } else {
            solve2<4>(S11b, cp, bp);
          }
          mm<2, 2, 4>(S21b, bp, bm);
#pragma unroll
          for (int i = 0; i < 8; ++i) Ebnext[i] = flushd(bm[i]);
        }
      }
    } else {  // CASE_LEFT
      cd am[8];
      mm<2, 2, 4>(SLs22, Fm, am);  // a+ = 0
      if (o == 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) Ef[i] = mk(0.0);
      } else {
        cd bm[8], bp[8], SLo12[4], SLo22[4];
        st.get4(SL_X0, SLo12);
        st.get4(SL_X1, SLo22);
        if (ill_conditioned(SLo22)) {
#pragma unroll
          for (int i = 0; i < 8; ++i) bm[i] = mk(0.0);
        } else {
          solve2<4>(SLo22, am, bm);
        }
        mm<2, 2, 4>(SLo12, bm, bp);
#pragma unroll
        for (int i = 0; i < 8; ++i) Ef[i] = flushd(bp[i]);
      }
      if (o + 1 == s) {
#pragma unroll
        for (int i = 0; i < 8; ++i) Ebnext[i] = Fm[i];
      } else {
        cd bm[8], SLo1_22[4];
        st.get4(SL_X2, SLo1_22);
        if (ill_conditioned(SLo1_22)) {
#pragma unroll
          for (int i = 0; i < 8; ++i) bm[i] = mk(0.0);
        } else {
          solve2<4>(SLo1_22, am, bm);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) Ebnext[i] = flushd(bm[i]);
      }
    }
    cd Eb[8];
    if (CASE == CASE_RIGHT && o == N - 1) {
#pragma unroll
      for (int i = 0; i < 8; ++i) Eb[i] = mk(0.0);
    } else {
      // re-reference: Eb = flush(S21(Sint_o) (Ef .* phf_o) + S22(Sint_o) Eb_next)
      const double dob = Lo.d;
      const cd ph[2] = {expi(Qo[0], dob), expi(Qo[1], dob)};
      cd far[8], t1[8], t2[8];
#pragma unroll
      for (int j = 0; j < 4; ++j) { far[2 * j] = ph[0] * Ef[2 * j]; far[2 * j + 1] = ph[1] * Ef[2 * j + 1]; }
      cd Sint21[4], Sint22[4];
      st.get4(SL_SI21, Sint21);
      st.get4(SL_SI22, Sint22);
      mm<2, 2, 4>(Sint21, far, t1);
      mm<2, 2, 4>(Sint22, Ebnext, t2);
#pragma unroll
      for (int i = 0; i < 8; ++i) Eb[i] = flushd(t1[i] + t2[i]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      T[4 * j] = Ef[2 * j]; T[4 * j + 1] = Ef[2 * j + 1];
      T[4 * j + 2] = Eb[2 * j]; T[4 * j + 3] = Eb[2 * j + 1];
    }
  }
  bool Tfin = true;
#pragma unroll
  for (int i = 0; i < 16; ++i) Tfin = Tfin && isfin(T[i]);

  // ---- node validity (mod_CavityTransfer.jl:226-251) ------------------------
  // a fast solve that met a singular interface is invalid until its redo
  sing = sing || redo;
  const bool valid = ok && !sing && seeds_fin;
  const bool Tvalid = valid && Tfin;
  cd basis[16];
  obs_basis(P.L[o], w, kx, ky, Vo, Qo, basis);
  const cd z0 = mk(0.0);
  const int64_t nn = P.nn;
  // the flags, the mark and the cache outputs come from the group of depth
  // pairs that starts at pair 0 (every group solves the node identically)
  const bool lead = pair0 == 0;
  if (lead && P.flags) P.flags[k] = (unsigned char)((valid ? 1 : 0) | (Tvalid ? 2 : 0) | (sing ? 4 : 0) | (ok ? 0 : 8));
  if (!ROBUST && __builtin_expect(redo, 0) && lead && P.redo) redo_mark(P.redo, k);
  if (lead && P.Qsrc) {
#pragma unroll
    for (int i = 0; i < 4; ++i) P.Qsrc[i * nn + k] = sel(valid, Qs[i], z0);
  }
  if (lead && P.sf) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      P.sf[i * nn + k] = sel(valid, sif[i], z0);
      P.sb[i * nn + k] = sel(valid, -sib[i], z0);
    }
  }
  if (lead && P.T) {
#pragma unroll
    for (int i = 0; i < 16; ++i) P.T[i * nn + k] = sel(Tvalid, T[i], z0);
  }
  if (lead && P.basis) {
#pragma unroll
    for (int i = 0; i < 16; ++i) P.basis[i * nn + k] = sel(valid, basis[i], z0);
  }

  // ---- spectral dyadic g~ for each (zsrc, zobs) pair -----------------------
  // an isotropic closed-form source layer in slot order: kze = kzo, so the two
  // modes of a side share their phases
  const bool iso_pair = phase_pairs && Ls.ek == 0;
  if (P.ghat) {
    const double d_src = Ls.d, d_obs = P.L[o].d;
    for (int pr = pair0; pr < pair1; ++pr) {
      const double zsl = P.zsrc[pr], zob = P.zobs[pr];
      const double zpos = s == 0 ? d_src - zsl : zsl;
      cd Af[6], Ab[6], Afd[6], Abd[6];  // [m + 2q]
#pragma unroll
      for (int i = 0; i < 6; ++i) { Af[i] = z0; Ab[i] = z0; Afd[i] = z0; Abd[i] = z0; }
      // term classes (_term_amplitudes, mod_ClosedFormField.jl:143-160): the
      // seeds feeding the forward / backward amplitudes, and the direct term
      const int terms = P.terms;
      const double dz = P.dz;
      if (Tvalid && terms != TERMS_DIRECT) {
        cd Pf[2], Pb[2];
        Pf[0] = expi(Qs[0], d_src - zpos);
        Pb[0] = expi(-Qs[2], zpos);
        if (iso_pair) {  // an isotropic closed-form layer: the TM phases are the TE ones
          Pf[1] = Pf[0];
          Pb[1] = Pb[0];
        } else {
          Pf[1] = expi(Qs[1], d_src - zpos);
          Pb[1] = expi(-Qs[3], zpos);
        }
        if (dz > 0.0) {  // box seeds D.Psi (_seed_phases)
#pragma unroll
          for (int m = 0; m < 2; ++m) {
            Pf[m] = boxphase(Qs[m], d_src - zpos, dz);
            Pb[m] = boxphase(Qs[2 + m], -zpos, dz);
          }
        }
        // seeds=:f keeps the forward seed columns (0, 1), seeds=:b the backward
        // (2, 3): weights of the two seed halves in the forward (rows 0-1)
        // and backward (rows 2-3) amplitudes, as arithmetic 0/1 weights.
        // Selects between the partial sums give the same results (every GPU
        // parity test passes with them) but made the headline node solve 1%
        // slower.
        const double wFf = terms == TERMS_HANKEL ? 0.0 : 1.0, wFb = terms == TERMS_TOEPLITZ ? 0.0 : 1.0;
        const double wBf = terms == TERMS_TOEPLITZ ? 0.0 : 1.0, wBb = terms == TERMS_HANKEL ? 0.0 : 1.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const cd u[4] = {sif[2 * q] * Pf[0], sif[2 * q + 1] * Pf[1], -(sib[2 * q] * Pb[0]), -(sib[2 * q + 1] * Pb[1])};
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const cd vf = fmac(T[i] * u[0], T[i + 4], u[1]);
            const cd vb = fmac(T[i + 8] * u[2], T[i + 12], u[3]);
            const cd v = rmul(i < 2 ? wFf : wBf, vf) + rmul(i < 2 ? wFb : wBb, vb);
            if (i < 2) Af[i + 2 * q] = v;
            else Ab[i - 2 + 2 * q] = v;
          }
        }
      }
      bool inside = false;
      if ((CASE == CASE_SAME || CASE == CASE_TOP) && valid &&
          (terms == TERMS_ALL || terms == TERMS_DIRECT || terms == TERMS_TOEPLITZ)) {
        const double zeta = zob - zpos;
        if (dz > 0.0 && fabs(zeta) < 0.5 * dz) {
          // inside the box's own slab: the face ramps (interior_slab_at)
          inside = true;
          const double a = fmin(fmax(zob - (zpos - 0.5 * dz), 0.0), dz);
          const double b = fmin(fmax((zpos + 0.5 * dz) - zob, 0.0), dz);
#pragma unroll
          for (int m = 0; m < 2; ++m) {
            const cd rf = ramp(Qs[m], a, dz), rb = ramp(-Qs[2 + m], b, dz);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              Afd[m + 2 * q] = sif[m + 2 * q] * rf;
              Abd[m + 2 * q] = -(sib[m + 2 * q] * rb);
            }
          }
        } else {
          // point: strict sign of zeta; box: |zeta| >= dz/2 (_direct_branch)
          const bool fwd = dz > 0.0 ? zeta >= 0.5 * dz : zeta > 0;
          const bool bwd = dz > 0.0 ? zeta <= -0.5 * dz : zeta < 0;
          if (fwd || bwd) {
            cd e[2];
#pragma unroll
            for (int m = 0; m < 2; ++m) {
              const cd Qd = fwd ? Qs[m] : Qs[2 + m];
              if (m == 1 && iso_pair && dz == 0.0) e[1] = e[0];
              else e[m] = dz > 0.0 ? boxphase(Qd, zeta, dz) : phase(Qd, zeta);
            }
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
              for (int q = 0; q < 3; ++q) {
                if (fwd) Afd[m + 2 * q] = sif[m + 2 * q] * e[m];
                else Abd[m + 2 * q] = -(sib[m + 2 * q] * e[m]);
              }
          }
        }
      }
      cd g[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) g[i] = z0;
      cd pfv[2], pbv[2];
      pfv[0] = phase(basis[12], zob);
      pbv[0] = phase(basis[14], zob - d_obs);
      if ((CASE == CASE_SAME || CASE == CASE_TOP) && iso_pair) {
        pfv[1] = pfv[0];
        pbv[1] = pbv[0];
      } else {
        pfv[1] = phase(basis[13], zob);
        pbv[1] = phase(basis[15], zob - d_obs);
      }
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const cd pf = pfv[m], pb = pbv[m];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const cd af = fin0(Af[m + 2 * q] * pf + Afd[m + 2 * q]);
          const cd ab = fin0(Ab[m + 2 * q] * pb + Abd[m + 2 * q]);
          g[0 + 3 * q] = g[0 + 3 * q] + (basis[0 + m] * af + basis[6 + m] * ab);
          g[1 + 3 * q] = g[1 + 3 * q] + (basis[2 + m] * af + basis[8 + m] * ab);
          g[2 + 3 * q] = g[2 + 3 * q] + (basis[4 + m] * af + basis[10 + m] * ab);
        }
      }
      // the local Ez inside the slab (ez_local_density): -p_z / (eps_zz dz)
      if (inside && P.ez_local) g[8] = g[8] - mk(1.0) / rmul(dz, Ls.eps[8]);
      cd *gp = P.ghat + (int64_t)pr * 9 * nn;
#pragma unroll
      for (int i = 0; i < 9; ++i) gp[i * nn + k] = sel(valid, g[i], z0);
    }
  }
}

// The fast node solve over every node (or every listed row).  A node whose
// solve met an exactly singular interface is marked in P.redo.  With
// gridDim.y = G > 1 the depth pairs are split into G contiguous groups, one
// per blockIdx.y: each group solves the node again and writes its pairs' g~,
// which fills the GPU when the nodes alone would not.
template <int CASE, bool GENERAL>
__global__ __launch_bounds__(64) void node_kernel(NodeParams P) {
  __shared__ cd lds[LdsSlots<CASE>::n * 64];
  const LaneMat st{lds, (int)threadIdx.x};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t n = P.n_rows ? min((int64_t)*P.n_rows, P.nn) : P.nn;
  const int ppg = (P.npairs + (int)gridDim.y - 1) / (int)gridDim.y;
  const int pair0 = (int)blockIdx.y * ppg, pair1 = min(P.npairs, pair0 + ppg);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    bool redo = false;
    solve_node<CASE, GENERAL, false>(P, k, st, redo, pair0, pair1);
  }
}

// The nodes (rows) that node_kernel marked, solved again with the reference's
// pivoted-QR fallback at the singular interfaces (mod_ScatteringMatrix.jl:35-44).
// Nothing to do, the usual case, costs one load per block.
template <int CASE, bool GENERAL>
__global__ __launch_bounds__(64) void node_kernel_redo(NodeParams P) {
  int *list = P.redo;
  if (!redo_any(list)) return;
  __shared__ cd lds[LdsSlots<CASE>::n * 64];
  const LaneMat st{lds, (int)threadIdx.x};
  NodeParams Q = P;
  Q.rows = nullptr;
  Q.n_rows = nullptr;
  for (int64_t k = blockIdx.x * 64 + threadIdx.x; k < P.nn; k += (int64_t)gridDim.x * 64) {
    if (!redo_take(list, k)) continue;
    bool again = false;
    solve_node<CASE, GENERAL, true>(Q, k, st, again, 0, Q.npairs);
  }
  redo_finish(list);
}

}  // namespace ytf
