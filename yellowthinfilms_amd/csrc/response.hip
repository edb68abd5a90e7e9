// response.hip -- the cavity response of a source layer at batches of
// in-plane wavevectors, FP64 on gfx950: the scan function of the real-axis
// pole search (_cavity_response, mod_KSpaceFeatures.jl:109-127, used by
// real_axis_poles :151-214).
//
// Per node (one lane): the Berreman modes of every layer with continuity
// matching and the interface S-matrices (as in batch.hip), then, for the
// source layer s (1-based),
//   RL = SLeft[s][1:2, 3:4]              (prefix Sprop[1] * Sint[1] * ... * Sint[s-1]),
//   RR = (Sint[s] * SRight[s+1])[3:4, 1:2]
//      = (Sint[s] * Sprop[s+1] * Sint[s+1] * ... * Sint[N-1])[3:4, 1:2],
// "*" the Redheffer star product.  The reference forms SRight right to left;
// here both products run left to right in the same pass over the layers
// (the star product is associative), and the prefix stops at the source.
//   r = max(sigma(RL), sigma(RR)), and for s > 1 also sigma(inv(I - PhiF RL PhiB RR)),
//   PhiF = e^{i Qf d_s}, PhiB = e^{-i Qb d_s},
// with sigma the largest singular value of a 2x2 (_σmax, :91-95).  A failed
// mode solve, an exactly singular interface solve (the reference falls back to
// a pivoted QR there) or a non-finite r give 0.
#include <hip/hip_runtime.h>

#include "node_core.hpp"

namespace ytf {

__device__ __forceinline__ double sigma_max2(cd a11, cd a21, cd a12, cd a22) {
  // G = A'A
  const double g11 = abs2(a11) + abs2(a21), g22 = abs2(a12) + abs2(a22);
  const cd g12 = conj(a11) * a12 + conj(a21) * a22;
  const double t = g11 + g22, dt = g11 * g22 - abs2(g12);
  return sqrt(fmax((t + sqrt(fmax(t * t - 4.0 * dt, 0.0))) * 0.5, 0.0));
}

template <bool GENERAL>
__device__ double response_node(const LayerDev *L, int nl, int s, double w, double kx, double ky, const LaneMat &Vp) {
  cd p[4], acc[16], R[16], RL[4] = {}, Qs[4];
  bool ok;
  {
    cd V[16], Q[4];
    ok = modes_layer<GENERAL>(L[0], w, kx, ky, V, Q);
#pragma unroll
    for (int i = 0; i < 16; ++i) Vp.set(i, V[i]);
#pragma unroll
    for (int i = 0; i < 4; ++i) Qs[i] = Q[i];
    prop4(Q, L[0].d, p);
  }
  bool singular = false;
  for (int j = 1; j < nl; ++j) {  // interface j between layers j-1 and j (0-based) = Sint[j]
    cd Q[4], Lm[16], Si[16];
    {
      cd V[16];
      ok = modes_layer<GENERAL>(L[j], w, kx, ky, V, Q) && ok;
      match_side_lds(Vp, V, Q, 0);
      match_side_lds(Vp, V, Q, 2);
      // _S_Interface_Matrix: [-Vf_B  Vb_A] \ [-Vf_A  Vb_B]
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        Lm[i] = -V[i];
        Lm[8 + i] = Vp.get(8 + i);
        Si[i] = -Vp.get(i);
        Si[8 + i] = V[8 + i];
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) Vp.set(i, V[i]);
    }
    singular = !lu_solve4<4>(Lm, Si) || singular;
    if (j < s) {  // prefix up to SLeft[s] = Sprop[1] * Sint[1] * ... * Sint[s-1]
      cd St[16];
      if (j == 1) redheffer_ldiag(p, Si, St);
      else redheffer(acc, Si, St);
      if (j == s - 1) {
        RL[0] = St[8];   // (1, 3)
        RL[1] = St[9];   // (2, 3)
        RL[2] = St[12];  // (1, 4)
        RL[3] = St[13];  // (2, 4)
      } else {
        cd pj[4];
        prop4(Q, L[j].d, pj);
        redheffer_rdiag(St, pj, acc);
      }
    }
    if (j == s) {
#pragma unroll
      for (int i = 0; i < 16; ++i) R[i] = Si[i];
    } else if (j > s) {  // R = (R * Sprop[j - 1]) * Sint[j]
      cd X[16];
      redheffer_rdiag(R, p, X);
      redheffer(X, Si, R);
    }
    if (j == s - 1) {
#pragma unroll
      for (int i = 0; i < 4; ++i) Qs[i] = Q[i];
    }
    prop4(Q, L[j].d, p);  // Sprop of layer j, for the next interface
  }
  // RR = R[3:4, 1:2]
  const cd r11 = R[2], r21 = R[3], r12 = R[6], r22 = R[7];
  double r = fmax(sigma_max2(RL[0], RL[1], RL[2], RL[3]), sigma_max2(r11, r21, r12, r22));
  if (s > 1) {
    const double d = L[s - 1].d;
    const cd f1 = expi(Qs[0], d), f2 = expi(Qs[1], d);     // PhiF
    const cd b1 = expi(-Qs[2], d), b2 = expi(-Qs[3], d);   // PhiB = e^{i Qb (-d)}
    // A = PhiF RL PhiB RR; M = I - A
    const cd l11 = f1 * RL[0] * b1, l21 = f2 * RL[1] * b1, l12 = f1 * RL[2] * b2, l22 = f2 * RL[3] * b2;
    const cd a11 = l11 * r11 + l12 * r21, a21 = l21 * r11 + l22 * r21;
    const cd a12 = l11 * r12 + l12 * r22, a22 = l21 * r12 + l22 * r22;
    const cd m11 = mk(1.0) - a11, m21 = -a21, m12 = -a12, m22 = mk(1.0) - a22;
    const cd det = m11 * m22 - m12 * m21;
    r = fmax(r, sigma_max2(m22 / det, -m21 / det, -m12 / det, m11 / det));
  }
  return (ok && !singular && isfinite(r)) ? r : 0.0;
}

struct ResponseArgs {
  const LayerDev *L;
  int nl, s;
  double omega;
  const double *kx, *ky;
  int64_t n;
  double *out;
};

template <bool GENERAL>
__global__ __launch_bounds__(64) void k_cavity_response(ResponseArgs A) {
  __shared__ cd vprev[16 * 64];
  const LaneMat Vp{vprev, (int)threadIdx.x};
  for (int64_t g = blockIdx.x * (int64_t)64 + threadIdx.x; g < A.n; g += (int64_t)gridDim.x * 64)
    A.out[g] = response_node<GENERAL>(A.L, A.nl, A.s, A.omega, A.kx[g], A.ky[g], Vp);
}

hipError_t launch_cavity_response(const LayerDev *L, int nl, int s, double omega, const double *kx, const double *ky,
                                  int64_t n, double *out, bool general, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  ResponseArgs A{L, nl, s, omega, kx, ky, n, out};
  int64_t blocks = (n + 63) / 64;
  if (blocks > 1 << 20) blocks = 1 << 20;
  if (general) hipLaunchKernelGGL(k_cavity_response<true>, dim3((unsigned)blocks), dim3(64), 0, st, A);
  else hipLaunchKernelGGL(k_cavity_response<false>, dim3((unsigned)blocks), dim3(64), 0, st, A);
  return hipGetLastError();
}

}  // namespace ytf
