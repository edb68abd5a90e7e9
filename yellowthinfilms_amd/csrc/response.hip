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
// with sigma the largest singular value of a 2x2 (_σmax, :91-95).  An exactly
// singular interface solve takes the reference's pivoted-QR fallback; a failed
// mode solve or a non-finite r give 0.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "node_core.hpp"

namespace ytf {

__device__ __forceinline__ double sigma_max2(cd a11, cd a21, cd a12, cd a22) {
  // G = A'A
  const double g11 = abs2(a11) + abs2(a21), g22 = abs2(a12) + abs2(a22);
  const cd g12 = conj(a11) * a12 + conj(a21) * a22;
  const double t = g11 + g22, dt = g11 * g22 - abs2(g12);
  return sqrt(fmax((t + sqrt(fmax(t * t - 4.0 * dt, 0.0))) * 0.5, 0.0));
}

// ROBUST = false: an interface LU that meets an exactly zero pivot only sets
// `redo` (see response), and the node is evaluated again with ROBUST = true,
// which takes the reference's pivoted-QR fallback (mod_ScatteringMatrix.jl:35-44).
template <bool GENERAL, bool ROBUST>
__device__ double response_node(const LayerDev *L, int nl, int s, double w, double kx, double ky, const LaneMat &Vp,
                                bool &redo) {
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
  uint64_t bits = 0;
  for (int j = 1; j < nl; ++j) {  // interface j between layers j-1 and j (0-based) = Sint[j]
    cd Q[4], Lm[16], Si[16];
    {
      cd V[16];
      ok = modes_layer<GENERAL>(L[j], w, kx, ky, V, Q) && ok;
      const bool sf = match_side_lds(Vp, V, Q, 0);
      const bool sb = match_side_lds(Vp, V, Q, 2);
      bits |= ((uint64_t)sf << (2 * j)) | ((uint64_t)sb << (2 * j + 1));
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
    const bool lu_ok = lu_solve4<4>(Lm, Si);
    if (!ROBUST) redo = redo || !lu_ok;
    else if (!lu_ok) {
      // SingularException: layer j-1 re-solved in its matched order, layer j
      // from LDS, and the pivoted-QR fallback
      cd Sf[16];
      interface_fallback<GENERAL>(&L[j - 1], w, kx, ky, (bits >> (2 * (j - 1))) & 1, (bits >> (2 * (j - 1) + 1)) & 1,
                                  Vp.p + Vp.lane, false, Sf);
#pragma unroll
      for (int i = 0; i < 16; ++i) Si[i] = Sf[i];
    }
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
  return (ok && !redo && isfinite(r)) ? r : 0.0;
}

struct ResponseArgs {
  const LayerDev *L;
  int nl, s;
  double omega;
  const double *kx, *ky;
  int64_t n;
  double *out;
  int *redo;  // redo marks (two words, a byte per node)
};

// The scan: a node whose interface solve was exactly singular gets 0 here and
// is marked for the redo launch.
template <bool GENERAL>
__global__ __launch_bounds__(64) void k_cavity_response(ResponseArgs A) {
  __shared__ cd vprev[16 * 64];
  const LaneMat Vp{vprev, (int)threadIdx.x};
  for (int64_t g = blockIdx.x * (int64_t)64 + threadIdx.x; g < A.n; g += (int64_t)gridDim.x * 64) {
    bool redo = false;
    const double r = response_node<GENERAL, false>(A.L, A.nl, A.s, A.omega, A.kx[g], A.ky[g], Vp, redo);
    A.out[g] = r;
    if (__builtin_expect(redo, 0)) redo_mark(A.redo, g);
  }
}

// The marked nodes again with the reference's pivoted-QR fallback at the
// singular interfaces (mod_ScatteringMatrix.jl:35-44); marks cleared.
template <bool GENERAL>
__global__ __launch_bounds__(64) void k_cavity_response_redo(ResponseArgs A) {
  if (!redo_any(A.redo)) return;
  __shared__ cd vprev[16 * 64];
  const LaneMat Vp{vprev, (int)threadIdx.x};
  for (int64_t g = blockIdx.x * 64 + threadIdx.x; g < A.n; g += (int64_t)gridDim.x * 64) {
    if (!redo_take(A.redo, g)) continue;
    bool again = false;
    A.out[g] = response_node<GENERAL, true>(A.L, A.nl, A.s, A.omega, A.kx[g], A.ky[g], Vp, again);
  }
  redo_finish(A.redo);
}

hipError_t launch_cavity_response(const LayerDev *L, int nl, int s, double omega, const double *kx, const double *ky,
                                  int64_t n, double *out, bool general, int *redo, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  ResponseArgs A{L, nl, s, omega, kx, ky, n, out, redo};
  int64_t blocks = (n + 63) / 64;
  if (blocks > 1 << 20) blocks = 1 << 20;
  if (general) hipLaunchKernelGGL(k_cavity_response<true>, dim3((unsigned)blocks), dim3(64), 0, st, A);
  else hipLaunchKernelGGL(k_cavity_response<false>, dim3((unsigned)blocks), dim3(64), 0, st, A);
  if (general) hipLaunchKernelGGL(k_cavity_response_redo<true>, dim3(kRedoBlocks), dim3(64), 0, st, A);
  else hipLaunchKernelGGL(k_cavity_response_redo<false>, dim3(kRedoBlocks), dim3(64), 0, st, A);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Prominent maxima of a real-axis scan (real_axis_poles, mod_KSpaceFeatures.jl:
// 151-214, step 2).  R[m][i] is the response along direction m at k = ks[i];
// Y[m][i] = y = log(R + 1e-300) (k_pole_log).

__device__ __forceinline__ double pole_lg(double r) { return log(r + 1e-300); }

// Y = log(R + 1e-300), elementwise.
__global__ void k_pole_log(int64_t m, const double *__restrict__ R, double *__restrict__ Y) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < m) Y[e] = pole_lg(R[e]);
}

// ymax[m] = max_i Y[m][i]: one block of 256 threads per direction.
__global__ void k_pole_rowmax(int64_t n, const double *__restrict__ Y, double *__restrict__ rmax) {
  __shared__ double part[256];
  const double *r = Y + blockIdx.x * n;
  double v = -HUGE_VAL;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) v = fmax(v, r[i]);
  part[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) rmax[blockIdx.x] = part[0];
}

// y[i] (1 <= i <= n - 2) is a local maximum when it is > its left and >= its
// right neighbour and >= max(log max_m R + log_floor_rel, log_floor_abs).  Its
// valleys: left, the samples after the last one above y[i] within W steps
// (from i - W if none), and its lowest (the last lowest); right, the samples
// before the first one above y[i] within W steps, and its lowest (the first
// lowest).  base = the higher valley, prom = y[i] - base; kept when prom >=
// log_prom: key m n + i and (base, prom, k of the left valley, k of the right
// valley, y[i]) are appended through *count (at most max_out).
//
// One wave per 64 samples of a direction finds the maxima among them, then
// walks each maximum's two sides together, 64 samples per step: a ballot
// finds the first sample above y[i] in the step, and a shuffle reduction
// the lowest sample before it (ties to the one met first walking out, as a
// one-sample-at-a-time walk keeps them).  Walks span up to W = ln 4 / scan
// step samples (~1,400), which one lane per maximum took serially.
__global__ __launch_bounds__(256) void k_pole_candidates(int nphi, int64_t n, int W, double log_prom,
                                                         double log_floor_rel, double log_floor_abs,
                                                         const double *__restrict__ Y,
                                                         const double *__restrict__ ymax,
                                                         const double *__restrict__ ks, int max_out,
                                                         int *__restrict__ count, int64_t *__restrict__ key,
                                                         double *__restrict__ val) {
  const int64_t gw = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t segs = (n + 63) / 64;
  const int m = (int)(gw / segs);
  if (m >= nphi) return;  // whole waves
  const int64_t i0 = (gw % segs) * 64, i = i0 + lane;
  const double *r = Y + (int64_t)m * n;
  const double ymin = fmax(ymax[m] + log_floor_rel, log_floor_abs);
  double y = 0.0;
  bool is_max = false;
  if (i >= 1 && i <= n - 2) {
    y = r[i];
    is_max = y > r[i - 1] && y >= r[i + 1] && y >= ymin;
  }
  unsigned long long mask = __ballot(is_max);
  while (mask) {
    const int bl = __ffsll((long long)mask) - 1;
    mask &= mask - 1;
    const int64_t im = i0 + bl;
    const double ym = __shfl(y, bl);
    double best[2];
    int64_t bpos[2];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      // side 0 walks left (j = start - lane), side 1 right (j = start + lane)
      const int64_t lim = side == 0 ? (im - W > 0 ? im - W : 0) : (im + W < n - 1 ? im + W : n - 1);
      double vmin = ym;
      int64_t pmin = im;
      for (int64_t j0 = side == 0 ? im - 1 : im + 1; side == 0 ? j0 >= lim : j0 <= lim; j0 += side == 0 ? -64 : 64) {
        const int64_t j = side == 0 ? j0 - lane : j0 + lane;
        const bool in = side == 0 ? j >= lim : j <= lim;
        const double v = in ? r[j] : 0.0;
        const unsigned long long above = __ballot(in && v > ym);
        const int first = above ? __ffsll((long long)above) - 1 : 64;
        double bv = (in && lane < first && !isnan(v)) ? v : HUGE_VAL;  // a NaN never becomes the lowest
        int64_t bj = j;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          const double ov = __shfl_xor(bv, off);
          const int64_t oj = __shfl_xor(bj, off);
          // the lower value; on a tie the sample met first (nearer im)
          const bool take = ov < bv || (ov == bv && (side == 0 ? oj > bj : oj < bj));
          bv = take ? ov : bv;
          bj = take ? oj : bj;
        }
        if (bv < vmin) {  // strictly lower: earlier steps keep their ties
          vmin = bv;
          pmin = bj;
        }
        if (above) break;
      }
      best[side] = vmin;
      bpos[side] = pmin;
    }
    const double base = fmax(best[0], best[1]), prom = ym - base;
    if (lane == 0 && prom >= log_prom) {
      const int o = atomicAdd(count, 1);
      if (o < max_out) {
        key[o] = (int64_t)m * n + im;
        val[5 * o] = base;
        val[5 * o + 1] = prom;
        val[5 * o + 2] = ks[bpos[0]];
        val[5 * o + 3] = ks[bpos[1]];
        val[5 * o + 4] = ym;
      }
    }
  }
}

hipError_t launch_pole_candidates(int nphi, int64_t n, int W, double log_prom, double log_floor_rel,
                                  double log_floor_abs, const double *R, const double *ks, double *Y, double *ymax,
                                  int max_out, int *count, int64_t *key, double *val, hipStream_t s) {
  if (nphi <= 0 || n <= 0) return hipSuccess;
  const int64_t tot = (int64_t)nphi * n;
  const unsigned blocks = (unsigned)((tot + 255) / 256);
  hipLaunchKernelGGL(k_pole_log, dim3(blocks), dim3(256), 0, s, tot, R, Y);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_pole_rowmax, dim3(nphi), dim3(256), 0, s, n, (const double *)Y, ymax);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int64_t waves = (int64_t)nphi * ((n + 63) / 64);  // one wave per 64 samples of a direction
  hipLaunchKernelGGL(k_pole_candidates, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, nphi, n, W, log_prom,
                     log_floor_rel, log_floor_abs, (const double *)Y, (const double *)ymax, ks, max_out, count, key,
                     val);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Location and half-width of the candidates (real_axis_poles step 3), every
// evaluation inside the kernel.  Chain 2c + side of candidate c = (phi, a, b,
// kl, kr, base, y) (cand[7c ..]):
//   golden-section search for the maximum of f(k) = response(k cos phi,
//   k sin phi) on [a, b] (at most 60 steps, until b - a <= 1e-12 b; both
//   chains of a candidate take the same steps), q = (a + b) / 2;
//   level = e^base + (max(f(q), e^y) - e^base) / sqrt 2;
//   bisection (at most 50 steps, until |hi - lo| <= 1e-12 q) between q and
//   the valley kl (side 0) or kr (side 1) for the crossing of the level.
// out[3c] = q, out[3c + 1 + side] = the crossing.
//
// Each chain takes 2^D lanes and advances D steps per evaluation round: the
// points a step visits depend only on the earlier steps' decisions, so the
// lanes evaluate the 2^D - 1 points of the decision tree of the next D steps
// (lane 0 the next point, lanes 1-2 the two possible points after it, lanes
// 3-6 the four after those; the last lane takes f(q) in the first bisection
// round), then every lane replays the D steps with the actual values.  The
// steps and the points they keep are those of one-step-at-a-time search.
// The kernel runs one wave per SIMD and every chain takes about the same
// time, so D = 3 is used when the chains fill at most one wave per SIMD and
// D = 2 otherwise (a second round of waves would cost more than the rounds
// the deeper tree saves).
// With `evals`, the number of responses the one-step search takes (the
// golden-section ones once per candidate) is added to *evals.
struct RefineArgs {
  const LayerDev *L;
  int nl, s;
  double omega;
  int64_t nc;
  const double *cand;
  double *out;
  unsigned long long *evals;
};

constexpr double GOLD = 0.61803398874989484820;  // (sqrt 5 - 1) / 2

// One golden-section step with decision `left` (the maximum is left of d).
__device__ __forceinline__ double gs_step(bool left, double &a, double &b, double &c, double &d) {
  if (left) {
    b = d;
    d = c;
    c = b - GOLD * (b - a);
    return c;
  }
  a = c;
  c = d;
  d = a + GOLD * (b - a);
  return d;
}

template <bool GENERAL, int D>
__global__ __launch_bounds__(64) void k_pole_refine(RefineArgs A) {
  constexpr int NL = 1 << D;  // lanes per chain: 2^D - 1 tree points and the f(q) lane
  __shared__ cd vprev[16 * 64];
  const LaneMat Vp{vprev, (int)threadIdx.x};
  const int64_t chain = (blockIdx.x * (int64_t)64 + threadIdx.x) >> D;
  const int l = threadIdx.x & (NL - 1), gb = threadIdx.x & ~(NL - 1);  // lane in the chain, its first lane
  if (chain >= 2 * A.nc) return;  // whole chains: 2 nc chains of NL lanes
  const int64_t c = chain >> 1;
  const int side = (int)(chain & 1);
  const double *cd7 = A.cand + 7 * c;
  const double phi = cd7[0], cphi = cos(phi), sphi = sin(phi);
  // (an exactly singular interface evaluates to 0 here: the refinement has no
  // redo pass; its points lie inside brackets of the scan)
  bool sing_seen = false;
  auto f = [&](double k) {
    return response_node<GENERAL, false>(A.L, A.nl, A.s, A.omega, k * cphi, k * sphi, Vp, sing_seen);
  };
  // lane l < NL - 1: node l of the tree (children of p: 2p + 1 on a true
  // decision, 2p + 2 on a false one); its level and the decisions on its path
  const int lev = l == 0 ? 0 : (l < 3 ? 1 : 2);
  const int a1 = lev == 2 ? (l - 1) / 2 : l;     // the level-1 ancestor
  const bool dec1 = lev >= 1 && (a1 & 1);        // level-1 decision
  const bool dec2 = lev >= 2 && (l & 1);         // level-2 decision
  unsigned long long ne = 0, ng = 0;
  // One evaluation per lane and round, at one call site.  Phases: 0 f(c) and
  // f(d) (lanes 0, 1); 1 golden-section rounds; 2 bisection rounds (lane 7
  // takes f(q) in the first); 3 done.
  double a = cd7[1], b = cd7[2];
  double cc = b - GOLD * (b - a), dd = a + GOLD * (b - a);
  double fc = 0.0, fd = 0.0, q = 0.0, lvl = 0.0, lo = 0.0, hi = 0.0;
  const double eb = exp(cd7[5]);
  int phase = 0, it = 0;
  bool first = true;
  while (phase < 3) {
    double x = 0.0;
    bool ev = false;
    if (phase == 0) {
      x = l == 0 ? cc : dd;
      ev = l < 2;
    } else if (phase == 1) {  // this lane's point: step 0 (known decision), then its path
      double ta = a, tb = b, tc = cc, td = dd;
      x = gs_step(fc >= fd, ta, tb, tc, td);
      if (lev >= 1) x = gs_step(dec1, ta, tb, tc, td);
      if (lev >= 2) x = gs_step(dec2, ta, tb, tc, td);
      ev = l < NL - 1;
    } else {
      double tlo = lo, thi = hi;
      x = (tlo + thi) / 2;
      if (lev >= 1) {
        (dec1 ? tlo : thi) = x;
        x = (tlo + thi) / 2;
      }
      if (lev >= 2) {
        (dec2 ? tlo : thi) = x;
        x = (tlo + thi) / 2;
      }
      const bool bis = it < 50 && !(fabs(hi - lo) <= 1e-12 * q);
      if (l == NL - 1) x = q;
      ev = l < NL - 1 ? bis : first;
    }
    const double v = ev ? f(x) : 0.0;
    if (phase == 0) {
      fc = __shfl(v, gb);
      fd = __shfl(v, gb + 1);
      ne += 2;
      phase = 1;
    } else if (phase == 1) {
      int node = 0;
      for (int j = 0; j < D; ++j) {
        if (!(it < 60 && !((b - a) <= 1e-12 * b))) break;
        const bool left = fc >= fd;
        if (j > 0) node = 2 * node + (left ? 1 : 2);
        const double fx = __shfl(v, gb + node);
        gs_step(left, a, b, cc, dd);
        if (left) {
          fd = fc;
          fc = fx;
        } else {
          fc = fd;
          fd = fx;
        }
        ++it;
        ++ne;
      }
    } else {
      if (first) {
        const double pk = fmax(__shfl(v, gb + NL - 1), exp(cd7[6]));
        lvl = eb + (pk - eb) / 1.41421356237309504880;
        first = false;
      }
      int node = 0;
      for (int j = 0; j < D; ++j) {
        if (!(it < 50 && !(fabs(hi - lo) <= 1e-12 * q))) break;
        const double mid = (lo + hi) / 2;
        const bool up = __shfl(v, gb + node) > lvl;
        if (up) lo = mid;
        else hi = mid;
        node = 2 * node + (up ? 1 : 2);
        ++it;
        ++ne;
      }
      if (!(it < 50 && !(fabs(hi - lo) <= 1e-12 * q))) phase = 3;
    }
    if (phase == 1 && !(it < 60 && !((b - a) <= 1e-12 * b))) {  // golden section done: start the bisection
      q = (a + b) / 2;
      ng = ne + 1;  // and f(q)
      ++ne;
      lo = q;
      hi = side ? cd7[4] : cd7[3];
      it = 0;
      phase = 2;
    }
  }
  if (l == 0) {
    if (side == 0) A.out[3 * c] = q;
    A.out[3 * c + 1 + side] = (lo + hi) / 2;
    if (A.evals) atomicAdd(A.evals, side == 0 ? ne : ne - ng);
  }
}

hipError_t launch_pole_refine(const LayerDev *L, int nl, int s, double omega, int64_t nc, const double *cand,
                              double *out, unsigned long long *evals, bool general, int n_cu, hipStream_t st) {
  if (nc <= 0) return hipSuccess;
  RefineArgs A{L, nl, s, omega, nc, cand, out, evals};
  // 2 chains per candidate; 8 lanes each if all waves fit one per SIMD, else 4
  const int64_t simds = 4 * (int64_t)std::max(n_cu, 1);
  const bool deep = (16 * nc + 63) / 64 <= simds;
  const unsigned blocks = (unsigned)(((deep ? 16 : 8) * nc + 63) / 64);
  if (deep) {
    if (general) hipLaunchKernelGGL((k_pole_refine<true, 3>), dim3(blocks), dim3(64), 0, st, A);
    else hipLaunchKernelGGL((k_pole_refine<false, 3>), dim3(blocks), dim3(64), 0, st, A);
  } else {
    if (general) hipLaunchKernelGGL((k_pole_refine<true, 2>), dim3(blocks), dim3(64), 0, st, A);
    else hipLaunchKernelGGL((k_pole_refine<false, 2>), dim3(blocks), dim3(64), 0, st, A);
  }
  return hipGetLastError();
}

}  // namespace ytf
