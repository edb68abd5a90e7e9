// capi_op.hip -- C-ABI of the 2D lateral sum and of the padded, z-structured
// Green operator (include/ytf.h: ytf_closed_form_G_2d, ytf_operator_*).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "ctx.hpp"

namespace {

using zc = std::complex<double>;

int upload(ytf_ctx *c, void *dst, const void *src, size_t bytes) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return YTF_OK;
}


// uniform step of a grid, or NaN if it is not uniform (_uniform_step,
// mod_GreenOperator.jl:383-386)
double uniform_step(const double *z, int n) {
  if (n <= 1) return 0.0;
  const double d = (z[n - 1] - z[0]) / (n - 1);
  for (int k = 0; k < n; ++k)
    if (std::fabs(z[k] - (z[0] + k * d)) > 1e-9 * std::max(std::fabs(d), 1.0)) return NAN;
  return d;
}

// A device allocation freed on scope exit (after the stream has drained).
struct DevBuf {
  ytf_ctx *c = nullptr;
  void *p = nullptr;
  ~DevBuf() {
    if (p) {
      hipStreamSynchronize(c->stream);
      hipFree(p);
    }
  }
};
int dev_alloc(ytf_ctx *c, DevBuf &b, size_t bytes) {
  b.c = c;
  HIPCHK(c, hipMalloc(&b.p, bytes < 256 ? 256 : bytes));
  return YTF_OK;
}

}  // namespace

struct OpBlock {
  int obs = 0;  // 1-based
  int kind = YTF_BLOCK_DENSE;
  std::vector<double> zo;
  cd *T = nullptr, *H = nullptr, *D = nullptr;  // [slot][9][Px*Py]
  int nT = 0, nH = 0, nDs = 0, nDo = 0;          // D slot = i + nDs * j
  cd *PsM = nullptr, *PoM = nullptr;             // fold / expansion matrices (skeleton)
  std::vector<int> sidx, oidx;                   // skeleton planes of each side (all if kept whole)
  std::vector<std::complex<double>> Ps, Po;      // interpolation matrices (N x k), empty if whole
  double serr = 0.0, oerr = 0.0;
  double build_ms = 0.0, resid = -1.0;
  int64_t bytes = 0;
};

struct ytf_operator {
  int src = 1;
  int Nx = 0, Ny = 0, Px = 0, Py = 0;
  double dx = 0, dy = 0;
  std::vector<double> zs;  // caller order
  bool flip = false;
  std::vector<OpBlock> blocks;
  // apply scratch: source spectra, skeleton-folded source, block output, skeleton output
  void *work = nullptr;
  size_t work_bytes = 0;
  // device copies of P and E for ytf_operator_apply_host
  void *stage = nullptr;
  size_t stage_bytes = 0;
};

namespace {

void free_block(OpBlock &b) {
  for (cd *p : {b.T, b.H, b.D, b.PsM, b.PoM})
    if (p) hipFree(p);
  b.T = b.H = b.D = b.PsM = b.PoM = nullptr;
}

// Inputs of one operator build shared by the plane fills.
struct Build {
  const ytf_layer *layers;
  int nl;
  double omega;
  int src;  // 1-based
  int64_t n;
  const double *kx, *ky;
  const cd *W;
  int Px, Py;
  double dx, dy;
  int lateral;
  double eps;
  double dz;     // box thickness (0: point sources)
  int ez_local;  // local Ez inside a box's slab
  // device: NUFFT points and weights, padded offsets
  double *xf = nullptr, *yf = nullptr;
  cd *wr = nullptr;
  double *xs = nullptr, *ys = nullptr;
  cd *twx = nullptr, *twy = nullptr;
  // phi-rotation fill: the n nodes are rot_nk k_rho x rot_nphi angles, k_rho
  // fastest; rot_kr: k_rho then rot_nk zeros (the phi = 0 nodes), rot_cs:
  // cos then sin of the angles (device)
  int64_t rot_nk = 0;
  int rot_nphi = 0;
  const double *rot_kr = nullptr, *rot_cs = nullptr;
};

// Every layer takes the closed-form branch, so the whole problem is invariant
// under rotations about z (_is_azimuthally_symmetric_stack,
// mod_ModeSolver.jl:426-434).
bool azimuthally_symmetric(const ytf_layer *L, int nl) {
  for (int l = 0; l < nl; ++l)
    if (!((L[l].eps_kind == YTF_EPS_ISOTROPIC || L[l].eps_kind == YTF_EPS_UNIAXIAL) && L[l].mu_kind == YTF_MU_ISOTROPIC))
      return false;
  return true;
}

// g~ [np][9][n] of every node of the build at np depth pairs: one node solve,
// or with the rotation fill one solve per k_rho at phi = 0 and every angle its
// rotation (build_cavity_transfer_2d's rotation branch, mod_CavityTransfer.jl:361-378).
int solve_nodes(ytf_ctx *c, const Build &B, int obs_layer, int np, const double *zs, const double *zo, cd *g,
                int terms) {
  if (B.rot_nphi == 0)
    return green_nodes_locked(c, B.layers, B.nl, B.omega, B.kx, B.ky, B.n, B.src, obs_layer, np, zs, zo, (double *)g,
                              nullptr, terms, B.dz, B.ez_local);
  DevBuf g0;
  int rc = dev_alloc(c, g0, sizeof(cd) * 9 * (size_t)np * (size_t)std::max<int64_t>(B.rot_nk, 1));
  if (rc) return rc;
  rc = green_nodes_locked(c, B.layers, B.nl, B.omega, B.rot_kr, B.rot_kr + B.rot_nk, B.rot_nk, B.src, obs_layer, np,
                          zs, zo, (double *)g0.p, nullptr, terms, B.dz, B.ez_local);
  if (rc) return rc;
  hipError_t e = timed(c, "rot_fill", [&] {
    return ytf::launch_rot_fill(B.rot_nk, B.rot_nphi, np, (const cd *)g0.p, B.rot_cs, g, c->stream);
  });
  if (e != hipSuccess) return fail(c, YTF_EHIP, "operator build: rotation fill: %s", hipGetErrorString(e));
  return YTF_OK;
}

struct Req {
  double zs, zo;  // local source and observation depths
  cd *dst;        // [9][Px*Py] stored plane
};

// Ĝ planes of the requests, one terms class: node solve with every request
// as a depth pair, lateral sum on the padded offsets, FFT, centring sign and
// 1/(Px Py) (_ghat_planes + _fill_by_source!, mod_GreenOperator.jl:124-155).
int fill_planes(ytf_ctx *c, const Build &B, int obs_layer, const std::vector<Req> &reqs, int terms) {
  if (reqs.empty()) return YTF_OK;
  const int64_t nn = (int64_t)B.Px * B.Py;
  const double per_pair = 9.0 * 16.0 * (double)(B.n + nn);
  const int PB = (int)std::max<double>(1.0, std::min<double>((double)reqs.size(), 3.0e9 / per_pair));
  DevBuf gb, fb;
  int rc = dev_alloc(c, gb, sizeof(cd) * 9 * (size_t)PB * (size_t)std::max<int64_t>(B.n, 1));
  if (rc) return rc;
  rc = dev_alloc(c, fb, sizeof(cd) * 9 * (size_t)PB * (size_t)nn);
  if (rc) return rc;
  cd *g = (cd *)gb.p, *f = (cd *)fb.p;
  std::vector<double> zs, zo;
  for (size_t r0 = 0; r0 < reqs.size(); r0 += PB) {
    const int np = (int)std::min<size_t>(PB, reqs.size() - r0);
    zs.resize(np);
    zo.resize(np);
    for (int p = 0; p < np; ++p) {
      zs[p] = reqs[r0 + p].zs;
      zo[p] = reqs[r0 + p].zo;
    }
    rc = solve_nodes(c, B, obs_layer, np, zs.data(), zo.data(), g, terms);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    if (B.lateral == YTF_LATERAL_NUFFT) {
      e = ytf::launch_lat_strengths(B.n, np, g, B.wr, g, 9, 1, c->stream);
      if (e != hipSuccess) return fail(c, YTF_EHIP, "operator build: strengths: %s", hipGetErrorString(e));
      rc = nufft2d1_locked(c, B.n, B.xf, B.yf, 9 * np, (const double *)g, B.Px, B.Py, B.eps, (double *)f);
      if (rc) return rc;
    } else {
      e = timed(c, "lateral_direct", [&] {
        return ytf::launch_lat_direct(B.n, B.kx, B.ky, B.W, np, g, B.xs, B.Px, B.ys, B.Py, f, 9, 1, c->stream);
      });
      if (e != hipSuccess) return fail(c, YTF_EHIP, "operator build: lateral sum: %s", hipGetErrorString(e));
    }
    for (int p = 0; p < np && e == hipSuccess; ++p) {
      cd *dst = reqs[r0 + p].dst;
      e = ytf::launch_opx(false, B.Px, f + (int64_t)p * 9 * nn, 9, 0, nn, B.Px, B.Px, B.Py, dst, B.Py, nn, 9, B.twx,
                          c->stream);
      if (e == hipSuccess)
        e = ytf::launch_opy(false, B.Py, dst, B.Px, nn, 9, B.Py, B.Py, 1.0 / (double)nn, 1, B.twy, c->stream);
    }
    if (e != hipSuccess) return fail(c, YTF_EHIP, "operator build: FFT of G: %s", hipGetErrorString(e));
  }
  return YTF_OK;
}

// Media rule of the structured form (_structured_by_media,
// mod_GreenOperator.jl:161-167, with _is_azimuthally_symmetric_stack,
// mod_ModeSolver.jl:426-434).
bool structured_by_media(const ytf_layer *L, int nl, int src) {
  const ytf_layer &s = L[src - 1];
  if (s.mu_kind != YTF_MU_ISOTROPIC) return false;
  if (s.eps_kind == YTF_EPS_ISOTROPIC) return true;
  if (s.eps_kind != YTF_EPS_UNIAXIAL) return false;
  for (int l = 0; l < nl; ++l)
    if (!((L[l].eps_kind == YTF_EPS_ISOTROPIC || L[l].eps_kind == YTF_EPS_UNIAXIAL) &&
          L[l].mu_kind == YTF_MU_ISOTROPIC))
      return false;
  return true;
}

// Numerical check of G_scattered(z, z') = T(z' - z) + H(z' + z) on six pixels
// (_check_structure, mod_GreenOperator.jl:174-182); NaN if the grids are too
// small to test.
int check_structure(ytf_ctx *c, const Build &B, const std::vector<double> &zs, const std::vector<double> &zo,
                    double *res) {
  const int Ns = (int)zs.size(), No = (int)zo.size();
  *res = NAN;
  if (Ns < 3 || No < 3) return YTF_OK;
  const int i0 = (Ns + 1) / 2 - 1, j0 = std::min((No + 1) / 2, No - 2) - 1;  // 0-based
  const int pi[4] = {i0, i0 + 1, i0 - 1, i0}, pj[4] = {j0, j0 + 1, j0 + 1, j0 + 2};
  double zsv[4], zov[4];
  for (int p = 0; p < 4; ++p) {
    zsv[p] = zs[pi[p]];
    zov[p] = zo[pj[p]];
  }
  const double xs[3] = {0.0, 0.17, 0.4}, ys[2] = {0.0, -0.11};
  DevBuf gb, ob, xb;
  int rc = dev_alloc(c, gb, sizeof(cd) * 36 * (size_t)std::max<int64_t>(B.n, 1));
  if (rc) return rc;
  rc = dev_alloc(c, ob, sizeof(cd) * 6 * 36);
  if (rc) return rc;
  rc = dev_alloc(c, xb, sizeof(double) * 5);
  if (rc) return rc;
  double xy[5] = {xs[0], xs[1], xs[2], ys[0], ys[1]};
  rc = upload(c, xb.p, xy, sizeof xy);
  if (rc) return rc;
  rc = solve_nodes(c, B, B.src, 4, zsv, zov, (cd *)gb.p, YTF_TERMS_SCATTERED);
  if (rc) return rc;
  const double *xd = (const double *)xb.p;
  hipError_t e = ytf::launch_lat_direct(B.n, B.kx, B.ky, B.W, 4, (const cd *)gb.p, xd, 3, xd + 3, 2, (cd *)ob.p, 9,
                                        1, c->stream);
  if (e != hipSuccess) return fail(c, YTF_EHIP, "structure check: %s", hipGetErrorString(e));
  std::vector<cd> h(6 * 36);
  HIPCHK(c, hipMemcpyAsync(h.data(), ob.p, sizeof(cd) * h.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // layout [p][e][y][x] with 6 pixels per (p, e)
  double nr = 0.0, nf = 0.0;
  for (int k = 0; k < 54; ++k) {
    const cd a = h[k], b = h[54 + k], d = h[108 + k], f = h[162 + k];
    const double rr = a.r - b.r - d.r + f.r, ri = a.i - b.i - d.i + f.i;
    nr += rr * rr + ri * ri;
    nf += a.r * a.r + a.i * a.i;
  }
  *res = std::sqrt(nr) / std::max(std::sqrt(nf), 1e-300);
  return YTF_OK;
}

int alloc_planes(ytf_ctx *c, cd **p, int nslots, int64_t nn, int64_t *bytes) {
  const size_t b = sizeof(cd) * 9 * (size_t)nslots * (size_t)nn;
  HIPCHK(c, hipMalloc((void **)p, b < 256 ? 256 : b));
  *bytes += (int64_t)b;
  return YTF_OK;
}

// Representative (i, j) pairs (0-based) of Toeplitz difference m = j - i and
// Hankel sum s = i + j (_toeplitz_pair, _hankel_pair, mod_GreenOperator.jl:389-402).
void toeplitz_pair(int m, int Ns, int No, int *i, int *j) {
  if (m >= 0) { *i = 0; *j = m; return; }
  if (Ns - 1 + m >= 0 && Ns - 1 + m <= No - 1) { *i = Ns - 1; *j = Ns - 1 + m; return; }
  *i = -m; *j = 0;
}
void hankel_pair(int s, int Ns, int No, int *i, int *j) {
  if (s <= No - 1) { *i = 0; *j = s; return; }
  if (s - (Ns - 1) >= 0 && s - (Ns - 1) <= No - 1) { *i = Ns - 1; *j = s - (Ns - 1); return; }
  *i = s - (No - 1); *j = No - 1;
}

// ---------------------------------------------------------------------------
// z skeletons (_zskeleton, z_skeletons, mod_GreenOperator.jl:186-359)
// ---------------------------------------------------------------------------
struct Skel {
  bool whole = true;
  int N = 0;
  std::vector<int> idx;  // skeleton planes, increasing
  std::vector<zc> P;     // N x k column-major, P[idx, :] = I
  double err = 0.0;
};

}  // namespace

// A set of z skeletons (z_skeletons, mod_GreenOperator.jl:335-359): the source
// side and one per observation layer, with the planes they were built on.
struct ytf_skeletons {
  int src = 1;
  std::vector<double> zs_int;  // source planes, increasing cavity-frame depth
  Skel src_sk;
  std::vector<int> obs;        // observation layers, ascending
  std::vector<std::vector<double>> zo;
  std::vector<Skel> obs_sk;
};

namespace {

// y += x over n doubles (the block-by-block sum of a grouped build).
__global__ void k_add_f64(double *__restrict__ y, const double *__restrict__ x, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] += x[i];
}

int add_planes(ytf_ctx *c, cd *dst, const cd *src, int64_t n_complex) {
  if (!dst || !src || n_complex <= 0) return YTF_OK;
  const int64_t n = 2 * n_complex;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_add_f64, dim3(grid), dim3(256), 0, c->stream, (double *)dst, (const double *)src, n);
  HIPCHK(c, hipGetLastError());
  return YTF_OK;
}

bool same_planes(const std::vector<double> &a, const double *b, int n) {
  if ((int)a.size() != n) return false;
  for (int i = 0; i < n; ++i)
    if (std::fabs(a[i] - b[i]) > 1e-12 * std::max(1.0, std::fabs(a[i]))) return false;
  return true;
}

// Column-pivoted Householder QR of the N x N matrix R in place (pivot: the
// largest remaining column norm, the first on ties, as LAPACK's geqp3 /
// qr(R, ColumnNorm())); p receives the column order.
void pivoted_qr(std::vector<zc> &R, int N, std::vector<int> &p) {
  p.resize(N);
  for (int j = 0; j < N; ++j) p[j] = j;
  for (int k = 0; k < N; ++k) {
    int jm = k;
    double best = -1.0;
    for (int j = k; j < N; ++j) {
      double s = 0.0;
      for (int i = k; i < N; ++i) s += std::norm(R[i + (size_t)N * j]);
      if (s > best) { best = s; jm = j; }
    }
    if (jm != k) {
      for (int i = 0; i < N; ++i) std::swap(R[i + (size_t)N * k], R[i + (size_t)N * jm]);
      std::swap(p[k], p[jm]);
    }
    double nx = 0.0;
    for (int i = k; i < N; ++i) nx += std::norm(R[i + (size_t)N * k]);
    if (nx < 1e-290) {
      // entries below 1e-145: dropped (the reflection's 2/vv would overflow)
      for (int i = k + 1; i < N; ++i) R[i + (size_t)N * k] = 0.0;
      continue;
    }
    nx = std::sqrt(nx);
    const zc alpha = R[k + (size_t)N * k];
    const double aa = std::abs(alpha);
    const zc ph = aa > 0.0 ? alpha / aa : zc(1.0, 0.0);
    const zc beta = -ph * nx;
    std::vector<zc> v(N - k);
    v[0] = alpha - beta;
    for (int i = k + 1; i < N; ++i) v[i - k] = R[i + (size_t)N * k];
    const double vv = 2.0 * nx * (nx + aa);
    for (int j = k + 1; j < N; ++j) {
      zc w = 0.0;
      for (int i = k; i < N; ++i) w += std::conj(v[i - k]) * R[i + (size_t)N * j];
      const zc g = 2.0 * w / vv;
      for (int i = k; i < N; ++i) R[i + (size_t)N * j] -= v[i - k] * g;
    }
    R[k + (size_t)N * k] = beta;
    for (int i = k + 1; i < N; ++i) R[i + (size_t)N * k] = 0.0;
  }
}

// The interpolative decomposition of rank k from the pivoted factor
// (interp in _zskeleton_serial_blas): P[p[q], q] = 1, the other rows
// Rt[1:k,1:k] \ Rt[1:k,k+1:N] transposed; skeleton planes sorted increasing.
void interp(const std::vector<zc> &Rt, const std::vector<int> &p, int N, int k, std::vector<int> &idx,
            std::vector<zc> &P) {
  std::vector<zc> Pk((size_t)N * k, 0.0);
  for (int q = 0; q < k; ++q) Pk[p[q] + (size_t)N * q] = 1.0;
  for (int c = k; c < N; ++c) {
    // T[:, c] = Rt[0:k, 0:k] \ Rt[0:k, c] (back substitution)
    std::vector<zc> t(k);
    for (int i = k - 1; i >= 0; --i) {
      zc s = Rt[i + (size_t)N * c];
      for (int m = i + 1; m < k; ++m) s -= Rt[i + (size_t)N * m] * t[m];
      // an exactly zero pivot (fewer independent rows than planes): zero weight
      const zc piv = Rt[i + (size_t)N * i];
      t[i] = piv == zc(0.0, 0.0) ? zc(0.0, 0.0) : s / piv;
    }
    for (int q = 0; q < k; ++q) Pk[p[c] + (size_t)N * q] = t[q];
  }
  std::vector<int> perm(k);
  for (int q = 0; q < k; ++q) perm[q] = q;
  std::sort(perm.begin(), perm.end(), [&](int a, int b) { return p[a] < p[b]; });
  idx.resize(k);
  P.assign((size_t)N * k, 0.0);
  for (int q = 0; q < k; ++q) {
    idx[q] = p[perm[q]];
    for (int i = 0; i < N; ++i) P[i + (size_t)N * q] = Pk[i + (size_t)N * perm[q]];
  }
}

// The skeleton of one side: family rows from the node solve's mode
// wavenumbers (Q: [4][n] device), tall-skinny QR on the device, pivoted QR
// and rank bisection with the error checked on every row on the device.
// Sides of more than 64 planes are kept whole.
int compute_skeleton(ytf_ctx *c, int64_t n, const cd *Q, const int fam[4], const double sig[2], const double cc[2],
                     const std::vector<double> &z, double tol, Skel *out) {
  const int N = (int)z.size();
  out->whole = true;
  out->N = N;
  out->idx.resize(N);
  for (int j = 0; j < N; ++j) out->idx[j] = j;
  out->P.assign((size_t)N * N, 0.0);
  for (int j = 0; j < N; ++j) out->P[j + (size_t)N * j] = 1.0;
  out->err = 0.0;
  if (N <= 1 || N > 64 || n == 0) return YTF_OK;
  const int64_t nrows = 4 * n;
  const int64_t nch = (nrows + 63) / 64;
  const int G = (int)std::min<int64_t>(256, nch);
  DevBuf zb, rb, eb;
  int rc = dev_alloc(c, zb, sizeof(double) * N);
  if (!rc) rc = dev_alloc(c, rb, sizeof(cd) * ((size_t)G + 1) * N * N);
  if (!rc) rc = dev_alloc(c, eb, sizeof(unsigned long long) + sizeof(int) * 64 + sizeof(cd) * 64 * 64 + 256);
  if (!rc) rc = upload(c, zb.p, z.data(), sizeof(double) * N);
  if (rc) return rc;
  ytf::SkFamRows F;
  F.Q = Q;
  F.n = n;
  for (int e = 0; e < 4; ++e) F.fam[e] = fam[e];
  for (int f = 0; f < 2; ++f) {
    F.sig[f] = sig[f];
    F.c[f] = cc[f];
  }
  F.z = (const double *)zb.p;
  cd *parts = (cd *)rb.p, *Rf = parts + (size_t)G * N * N;
  hipError_t e = ytf::launch_sk_tsqr(F, nullptr, 0, N, nrows, parts, G, c->stream);
  if (e == hipSuccess) e = ytf::launch_sk_tsqr(F, parts, 1, N, (int64_t)G * N, Rf, 1, c->stream);
  if (e != hipSuccess) return fail(c, YTF_EHIP, "z skeleton: QR: %s", hipGetErrorString(e));
  std::vector<zc> R((size_t)N * N);
  HIPCHK(c, hipMemcpyAsync(R.data(), Rf, sizeof(cd) * R.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int> p;
  pivoted_qr(R, N, p);
  unsigned long long *errd = (unsigned long long *)eb.p;
  int *idxd = (int *)((char *)eb.p + 256);
  cd *Pd = (cd *)((char *)eb.p + 256 + sizeof(int) * 64);
  auto maxerr = [&](const std::vector<int> &idx, const std::vector<zc> &P, double *err) -> int {
    const int k = (int)idx.size();
    std::vector<char> h(sizeof(int) * 64 + sizeof(cd) * (size_t)N * k);
    std::memcpy(h.data(), idx.data(), sizeof(int) * k);
    std::memcpy(h.data() + sizeof(int) * 64, P.data(), sizeof(cd) * (size_t)N * k);
    int r = upload(c, idxd, h.data(), h.size());
    if (r) return r;
    hipError_t e2 = ytf::launch_sk_err(F, nrows, N, k, idxd, Pd, errd, c->stream);
    if (e2 != hipSuccess) return fail(c, YTF_EHIP, "z skeleton: error check: %s", hipGetErrorString(e2));
    unsigned long long bits = 0;
    HIPCHK(c, hipMemcpyAsync(&bits, errd, sizeof bits, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(err, &bits, sizeof(double));
    return YTF_OK;
  };
  int lo = 0, hi = N;
  std::vector<int> idx;
  std::vector<zc> P;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    interp(R, p, N, mid, idx, P);
    double err = 0.0;
    rc = maxerr(idx, P, &err);
    if (rc) return rc;
    if (err <= tol) hi = mid;
    else lo = mid;
  }
  if (hi >= 0.8 * N) return YTF_OK;  // whole_frac
  interp(R, p, N, hi, idx, P);
  double err = 0.0;
  rc = maxerr(idx, P, &err);
  if (rc) return rc;
  out->whole = false;
  out->idx = idx;
  out->P = P;
  out->err = err;
  return YTF_OK;
}

// Dense plane pairs (_dense_planes, mod_GreenOperator.jl:568-574) on the
// source planes zs and observation planes zo (all of a side's planes, or its
// skeleton planes).
int fill_dense(ytf_ctx *c, const Build &B, OpBlock &b, const std::vector<double> &zs, int terms,
               const std::vector<double> *zo_sub = nullptr) {
  const int64_t nn = (int64_t)B.Px * B.Py;
  const std::vector<double> &zo = zo_sub ? *zo_sub : b.zo;
  b.nDs = (int)zs.size();
  b.nDo = (int)zo.size();
  int rc = alloc_planes(c, &b.D, b.nDs * b.nDo, nn, &b.bytes);
  if (rc) return rc;
  std::vector<Req> reqs;
  for (int j = 0; j < b.nDo; ++j)
    for (int i = 0; i < b.nDs; ++i) reqs.push_back({zs[i], zo[j], b.D + (int64_t)(i + b.nDs * j) * 9 * nn});
  return fill_planes(c, B, b.obs, reqs, terms);
}

// Source side (_src_family, mod_GreenOperator.jl:329-333): planes at
// cavity-frame depth zpos, Psi_f = e^{iQf (d - zpos)}, Psi_b = e^{-iQb zpos}; a
// source in layer 1 feeds through Psi_f only.  Q: [4][n] Qf1 Qf2 Qb1 Qb2.
int src_skeleton(ytf_ctx *c, const ytf_layer *layers, int src, int64_t n, const cd *Q,
                 const std::vector<double> &zs_int, double tol, Skel *out) {
  const double d = layers[src - 1].d;
  std::vector<double> zpos(zs_int);
  if (src == 1)
    for (auto &v : zpos) v = d - v;
  const int fam[4] = {0, 0, src != 1 ? 1 : -1, src != 1 ? 1 : -1};
  const double sig[2] = {-1.0, -1.0}, cc[2] = {d, 0.0};
  return compute_skeleton(c, n, Q, fam, sig, cc, zpos, tol, out);
}

// Observation side (_obs_family, :318-323): e^{iKZf z'}, e^{iKZb (z' - d)};
// layer 1 carries the downward wave only, the last layer the upward one only.
// Q: [4][n] KZf1 KZf2 KZb1 KZb2.
int obs_skeleton(ytf_ctx *c, const ytf_layer *layers, int nl, int obs, int64_t n, const cd *Q,
                 const std::vector<double> &zo, double tol, Skel *out) {
  const double d = layers[obs - 1].d;
  const int fam[4] = {obs != 1 ? 0 : -1, obs != 1 ? 0 : -1, obs != nl ? 1 : -1, obs != nl ? 1 : -1};
  const double sig[2] = {1.0, 1.0}, cc[2] = {0.0, -d};
  return compute_skeleton(c, n, Q, fam, sig, cc, zo, tol, out);
}

// Skeletons of the source side and of observation layer `obs` from one node
// solve's mode wavenumbers (_cache_skeletons, _src_family, _obs_family,
// mod_GreenOperator.jl:315-359).
int side_skeletons(ytf_ctx *c, const Build &B, int obs, const std::vector<double> &zs_int,
                   const std::vector<double> &zo, double tol, Skel *src_sk, bool need_src, Skel *obs_sk) {
  // the mode wavenumbers do not depend on the angle: with the rotation fill,
  // each k_rho once
  const bool rot = B.rot_nphi > 0;
  const int64_t n = rot ? B.rot_nk : B.n;
  const double *kx = rot ? B.rot_kr : B.kx, *ky = rot ? B.rot_kr + B.rot_nk : B.ky;
  DevBuf qb;
  int rc = dev_alloc(c, qb, sizeof(cd) * 20 * (size_t)std::max<int64_t>(n, 1));
  if (rc) return rc;
  ytf_node_cache nc;
  std::memset(&nc, 0, sizeof nc);
  nc.Qsrc_dev = (double *)qb.p;
  nc.basis_dev = (double *)((cd *)qb.p + 4 * n);
  rc = green_nodes_locked(c, B.layers, B.nl, B.omega, kx, ky, n, B.src, obs, 0, nullptr, nullptr, nullptr, &nc,
                          YTF_TERMS_ALL, 0.0, 1);
  if (rc) return rc;
  if (need_src) {
    rc = src_skeleton(c, B.layers, B.src, n, (const cd *)qb.p, zs_int, tol, src_sk);
    if (rc) return rc;
  }
  return obs_skeleton(c, B.layers, B.nl, obs, n, (const cd *)qb.p + 16 * n, zo, tol, obs_sk);
}

// The skeletons= of build_layer_operator: the given set's source side and its
// side for observation layer `obs`, which must have been built on these planes.
int given_skeletons(ytf_ctx *c, const ytf_skeletons *sk, int src, const std::vector<double> &zs_int, int obs,
                    const std::vector<double> &zo, Skel *src_sk, Skel *obs_sk) {
  if (sk->src != src || !same_planes(sk->zs_int, zs_int.data(), (int)zs_int.size()))
    return fail(c, YTF_EINVAL, "build_layer_operator: skeletons were built for other source planes");
  for (size_t l = 0; l < sk->obs.size(); ++l)
    if (sk->obs[l] == obs) {
      if (!same_planes(sk->zo[l], zo.data(), (int)zo.size()))
        return fail(c, YTF_EINVAL, "build_layer_operator: skeletons were built for other planes of layer %d", obs);
      *src_sk = sk->src_sk;
      *obs_sk = sk->obs_sk[l];
      return YTF_OK;
    }
  return fail(c, YTF_EINVAL, "build_layer_operator: skeletons have no side for observation layer %d", obs);
}

int upload_matrix(ytf_ctx *c, cd **dst, const std::vector<zc> &h, int64_t *bytes) {
  HIPCHK(c, hipMalloc((void **)dst, sizeof(cd) * std::max<size_t>(h.size(), 1)));
  *bytes += (int64_t)(sizeof(cd) * h.size());
  return upload(c, *dst, h.data(), sizeof(cd) * h.size());
}

// The dense part of a block without z structure, on skeleton planes where a
// side compresses (_planes_maybe_skeleton, mod_GreenOperator.jl:581-587).
int fill_maybe_skeleton(ytf_ctx *c, const Build &B, OpBlock &b, const std::vector<double> &zs_int, int terms,
                        const Skel *ss, const Skel *os) {
  if (!ss || !os || (ss->whole && os->whole)) return fill_dense(c, B, b, zs_int, terms);
  std::vector<double> zs, zo;
  for (int i : ss->idx) zs.push_back(zs_int[i]);
  for (int j : os->idx) zo.push_back(b.zo[j]);
  int rc = fill_dense(c, B, b, zs, terms, &zo);
  if (rc) return rc;
  b.sidx = ss->idx;
  b.oidx = os->idx;
  if (!ss->whole) b.Ps = ss->P;
  if (!os->whole) b.Po = os->P;
  b.serr = ss->err;
  b.oerr = os->err;
  if (!ss->whole) {
    rc = upload_matrix(c, &b.PsM, ss->P, &b.bytes);  // (Ns x Rs): fold M[i + Ns q] = Ps[i, q]
    if (rc) return rc;
  }
  if (!os->whole) {
    const int No = (int)b.zo.size(), Ro = (int)os->idx.size();
    std::vector<zc> t((size_t)Ro * No);
    for (int j = 0; j < No; ++j)
      for (int q = 0; q < Ro; ++q) t[q + (size_t)Ro * j] = os->P[j + (size_t)No * q];  // M[p + Ro j] = Po[j, p]
    rc = upload_matrix(c, &b.PoM, t, &b.bytes);
    if (rc) return rc;
  }
  return YTF_OK;
}

int fill_toeplitz(ytf_ctx *c, const Build &B, OpBlock &b, const std::vector<double> &zs, int terms) {
  const int64_t nn = (int64_t)B.Px * B.Py;
  const int Ns = (int)zs.size(), No = (int)b.zo.size();
  b.nT = Ns + No - 1;
  int rc = alloc_planes(c, &b.T, b.nT, nn, &b.bytes);
  if (rc) return rc;
  std::vector<Req> reqs;
  for (int m = -(Ns - 1); m <= No - 1; ++m) {
    int i, j;
    toeplitz_pair(m, Ns, No, &i, &j);
    reqs.push_back({zs[i], b.zo[j], b.T + (int64_t)(m + Ns - 1) * 9 * nn});
  }
  return fill_planes(c, B, b.obs, reqs, terms);
}

int fill_hankel(ytf_ctx *c, const Build &B, OpBlock &b, const std::vector<double> &zs) {
  const int64_t nn = (int64_t)B.Px * B.Py;
  const int Ns = (int)zs.size(), No = (int)b.zo.size();
  b.nH = Ns + No - 1;
  int rc = alloc_planes(c, &b.H, b.nH, nn, &b.bytes);
  if (rc) return rc;
  std::vector<Req> reqs;
  for (int s = 0; s <= Ns + No - 2; ++s) {
    int i, j;
    hankel_pair(s, Ns, No, &i, &j);
    reqs.push_back({zs[i], b.zo[j], b.H + (int64_t)s * 9 * nn});
  }
  return fill_planes(c, B, b.obs, reqs, YTF_TERMS_HANKEL);
}

// With one source or one observation plane every Hankel plane coincides with
// one Toeplitz plane (i + j = s and j - i = m pick the same pair), and with
// one of each the dense scattered plane does too: store their sums, so the
// apply reads one plane per pair.
int merge_planes(ytf_ctx *c, OpBlock &b, int Ns, int64_t nn) {
  const int No = (int)b.zo.size();
  const size_t pb = sizeof(cd) * 9 * (size_t)nn;
  if (b.H && b.T && (Ns == 1 || No == 1)) {
    for (int s = 0; s < b.nH; ++s) {
      const int i = Ns == 1 ? 0 : s, j = Ns == 1 ? s : 0;
      hipError_t e = ytf::launch_op_add(9 * nn, b.H + (int64_t)s * 9 * nn, b.T + (int64_t)(j - i + Ns - 1) * 9 * nn,
                                        c->stream);
      if (e != hipSuccess) return fail(c, YTF_EHIP, "operator build: merge: %s", hipGetErrorString(e));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(b.H));
    b.H = nullptr;
    b.bytes -= (int64_t)(pb * b.nH);
    b.nH = 0;
  }
  if (b.D && b.T && Ns == 1 && No == 1 && b.nT == 1 && !b.PsM && !b.PoM) {
    hipError_t e = ytf::launch_op_add(9 * nn, b.D, b.T, c->stream);
    if (e != hipSuccess) return fail(c, YTF_EHIP, "operator build: merge: %s", hipGetErrorString(e));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(b.D));
    b.D = nullptr;
    b.bytes -= (int64_t)pb;
    b.nDs = b.nDo = 0;
  }
  return YTF_OK;
}

// The block is one stored plane between one source and one observation plane.
bool single_plane(const OpBlock &b, int Ns) {
  return Ns == 1 && b.zo.size() == 1 && !b.H && !b.PsM && !b.PoM && ((b.T && !b.D && b.nT == 1) || (b.D && !b.T));
}

}  // namespace

extern "C" {

int ytf_polar_nodes(ytf_ctx *c, int64_t n_krho, const double *krho, int n_phi, const double *phi,
                    const double *wphi, double box_dx, double box_dy, double *kx, double *ky, double *w) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_polar_nodes: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n_krho < 0 || n_phi < 1 || n_krho > INT64_MAX / n_phi || !(box_dx >= 0.0) || !(box_dy >= 0.0))
    return fail(c, YTF_EINVAL, "ytf_polar_nodes: bad sizes or box");
  if (n_krho == 0) return YTF_OK;
  if (!krho || !phi || !wphi || !kx || !ky || !w) return fail(c, YTF_EINVAL, "ytf_polar_nodes: NULL buffer");
  HIPCHK(c, hipSetDevice(c->device));
  hipError_t e = ytf::launch_polar_nodes(n_krho, n_phi, krho, phi, wphi, box_dx, box_dy, kx, ky, (cd *)w, c->stream);
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_polar_nodes: launch failed: %s", hipGetErrorString(e));
  return YTF_OK;
}

int ytf_segment_lateral_sum(ytf_ctx *c, int64_t n_seg, int seg_len, const double *kx, const double *ky,
                            const double *w, int np, const double *ghat, const double *xp, const double *yp,
                            double *out) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_segment_lateral_sum: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n_seg < 0 || n_seg > INT32_MAX || seg_len < 1 || np < 1 || np > 65535)
    return fail(c, YTF_EINVAL, "ytf_segment_lateral_sum: bad sizes (n_seg=%lld, seg_len=%d, n_pairs=%d)",
                (long long)n_seg, seg_len, np);
  if (n_seg == 0) return YTF_OK;
  if (!kx || !ky || !w || !ghat || !xp || !yp || !out)
    return fail(c, YTF_EINVAL, "ytf_segment_lateral_sum: NULL buffer");
  HIPCHK(c, hipSetDevice(c->device));
  // up to LatPts::kMax points travel with the launch (no upload, so no wait
  // for earlier kernels); more go through the context's buffer
  ytf::LatPts pts{};
  const double *xd = nullptr, *yd = nullptr;
  if (np <= ytf::LatPts::kMax) {
    std::copy(xp, xp + np, pts.x);
    std::copy(yp, yp + np, pts.y);
  } else {
    int rc = ensure_buf(c, &c->aux, &c->aux_bytes, sizeof(double) * 2 * (size_t)np);
    if (rc) return rc;
    std::vector<double> h(xp, xp + np);
    h.insert(h.end(), yp, yp + np);
    rc = upload(c, c->aux, h.data(), sizeof(double) * h.size());
    if (rc) return rc;
    xd = (const double *)c->aux;
    yd = xd + np;
  }
  hipError_t e = timed(c, "lateral_segments", [&] {
    return ytf::launch_lat_segments(n_seg, seg_len, kx, ky, (const cd *)w, np, (const cd *)ghat, xd, yd, pts,
                                    (cd *)out, c->stream);
  });
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_segment_lateral_sum: launch failed: %s", hipGetErrorString(e));
  return YTF_OK;
}

int ytf_closed_form_G_2d(ytf_ctx *c, int64_t n, const double *kx, const double *ky, const double *w, int np,
                         const double *ghat, const double *xs, int nx, const double *ys, int ny, int lateral,
                         double eps, double *G) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_closed_form_G_2d: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (lateral != YTF_LATERAL_DIRECT && lateral != YTF_LATERAL_NUFFT)
    return fail(c, YTF_EINVAL, "closed_form_G: lateral must be :direct or :nufft");
  if (n < 0 || np < 1 || nx < 1 || ny < 1) return fail(c, YTF_EINVAL, "ytf_closed_form_G_2d: bad sizes");
  if (!G || !xs || !ys || (n > 0 && (!kx || !ky || !w || !ghat)))
    return fail(c, YTF_EINVAL, "ytf_closed_form_G_2d: NULL buffer");
  HIPCHK(c, hipSetDevice(c->device));
  if (lateral == YTF_LATERAL_DIRECT) {
    int rc = ensure_buf(c, &c->aux, &c->aux_bytes, sizeof(double) * (nx + ny));
    if (rc) return rc;
    std::vector<double> h(xs, xs + nx);
    h.insert(h.end(), ys, ys + ny);
    rc = upload(c, c->aux, h.data(), sizeof(double) * h.size());
    if (rc) return rc;
    const double *xd = (const double *)c->aux;
    hipError_t e = timed(c, "lateral_direct", [&] {
      return ytf::launch_lat_direct(n, kx, ky, (const cd *)w, np, (const cd *)ghat, xd, nx, xd + nx, ny, (cd *)G, 1,
                                    np, c->stream);
    });
    if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_closed_form_G_2d: launch failed: %s", hipGetErrorString(e));
    return YTF_OK;
  }
  if (nx > 2048 || ny > 2048 || nx < 2 || ny < 2)
    return fail(c, YTF_EINVAL, "closed_form_G: lateral=:nufft supports 2 x 2 to 2048 x 2048 pixels (got %d x %d)",
                nx, ny);
  const double sx = uniform_step(xs, nx), sy = uniform_step(ys, ny);
  if (std::isnan(sx) || std::isnan(sy))
    return fail(c, YTF_EINVAL, "closed_form_G: lateral=:nufft needs uniform xs and ys");
  // _nufft_nodes_2d with the centre pixel cx = 1 + fld(Nx, 2) (1-based) and
  // x0 = -xs[cx]: points kx px folded by rem2pi, ramp exp(i (kx xs[cx] + ky ys[cy]))
  const double px = nx > 1 ? xs[1] - xs[0] : 1.0, py = ny > 1 ? ys[1] - ys[0] : 1.0;
  const double x0 = xs[nx / 2], y0 = ys[ny / 2];
  const size_t nb = (size_t)std::max<int64_t>(n, 1);
  const size_t b_pts = align256(sizeof(double) * nb), b_w = align256(sizeof(cd) * nb),
               b_s = align256(sizeof(cd) * 9 * (size_t)np * nb);
  int rc = ensure_buf(c, &c->aux, &c->aux_bytes, 2 * b_pts + b_w + b_s);
  if (rc) return rc;
  char *a = (char *)c->aux;
  double *xf = (double *)a, *yf = (double *)(a + b_pts);
  cd *wr = (cd *)(a + 2 * b_pts), *sv = (cd *)(a + 2 * b_pts + b_w);
  hipError_t e = ytf::launch_lat_points(n, kx, ky, (const cd *)w, px, py, x0, y0, xf, yf, wr, c->stream);
  if (e == hipSuccess) e = ytf::launch_lat_strengths(n, np, (const cd *)ghat, wr, sv, 1, np, c->stream);
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_closed_form_G_2d: launch failed: %s", hipGetErrorString(e));
  return nufft2d1_locked(c, n, xf, yf, 9 * np, (const double *)sv, nx, ny, eps, G);
}

int ytf_operator_build(ytf_ctx *c, const ytf_layer *layers, int n_layers, double omega, int source_layer,
                       int64_t n_nodes, const double *kx_dev, const double *ky_dev, const double *w_dev,
                       const ytf_grid *grid, const double *z_src, int n_src, int n_obs_layers,
                       const int32_t *obs_layers, const int32_t *n_obs, const double *z_obs,
                       const ytf_operator_options *opt, uint32_t *warnings, ytf_operator **out) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_operator_build: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!out || !grid || !opt || !layers) return fail(c, YTF_EINVAL, "ytf_operator_build: NULL argument");
  *out = nullptr;
  if (warnings) *warnings = 0;
  if (opt->structure < YTF_STRUCTURE_AUTO || opt->structure > YTF_STRUCTURE_DENSE)
    return fail(c, YTF_EINVAL, "build_layer_operator: structure must be :auto, :structured or :dense");
  if (opt->lateral != YTF_LATERAL_DIRECT && opt->lateral != YTF_LATERAL_NUFFT)
    return fail(c, YTF_EINVAL, "closed_form_G: lateral must be :direct or :nufft");
  if (grid->Nx < 1 || grid->Ny < 1 || !ytf::fft_length_ok(2 * grid->Nx) || !ytf::fft_length_ok(2 * grid->Ny))
    return fail(c, YTF_EINVAL,
                "ytf_operator_build: grid Nx=%d, Ny=%d unsupported (powers of two up to 1024, any other size up to 512)",
                grid->Nx, grid->Ny);
  if (!(grid->dx > 0) || !(grid->dy > 0)) return fail(c, YTF_EINVAL, "ytf_operator_build: grid pitch must be > 0");
  if (n_src < 1 || !z_src) return fail(c, YTF_EINVAL, "build_layer_operator: need at least one source plane");
  if (n_obs_layers < 0 || (n_obs_layers > 0 && (!obs_layers || !n_obs)))
    return fail(c, YTF_EINVAL, "ytf_operator_build: bad observation layers");
  if (n_nodes < 0 || (n_nodes > 0 && (!kx_dev || !ky_dev || !w_dev)))
    return fail(c, YTF_EINVAL, "ytf_operator_build: bad nodes");
  if (source_layer < 1 || source_layer > n_layers - 1)
    return fail(c, YTF_EINVAL, "ytf_operator_build: source_layer=%d must be in 1:%d", source_layer, n_layers - 1);
  if (opt->rot_n_phi < 0 ||
      (opt->rot_n_phi > 0 && (opt->rot_n_krho < 0 || !opt->rot_krho || !opt->rot_phi ||
                              opt->rot_n_krho > INT64_MAX / opt->rot_n_phi ||
                              opt->rot_n_krho * (int64_t)opt->rot_n_phi != n_nodes)))
    return fail(c, YTF_EINVAL, "ytf_operator_build: the rotation fill needs n_nodes = rot_n_krho x rot_n_phi and both "
                               "axes");
  if (opt->rot_n_phi > 0 && !azimuthally_symmetric(layers, n_layers))
    return fail(c, YTF_EINVAL, "build_cavity_transfer_2d: rotation=true requires an azimuthally symmetric stack");
  const bool skeleton = opt->skeleton && opt->structure != YTF_STRUCTURE_DENSE;
  // the reference requires uniform z grids (errors otherwise)
  double ds = 0.0;
  if (opt->uniform_z) {
    ds = uniform_step(z_src, n_src);
    if (std::isnan(ds)) return fail(c, YTF_EINVAL, "build_layer_operator: z grids must be uniform");
  } else if (opt->structure != YTF_STRUCTURE_DENSE) {
    return fail(c, YTF_EINVAL, "ytf_operator_build: arbitrary depths need structure=:dense");
  }
  // layers sorted ascending, repeats rejected
  std::vector<std::pair<int, int>> order;  // (layer, position in the inputs)
  int64_t off = 0;
  std::vector<int64_t> zoff(n_obs_layers);
  for (int l = 0; l < n_obs_layers; ++l) {
    if (obs_layers[l] < 1 || obs_layers[l] > n_layers || n_obs[l] < 0)
      return fail(c, YTF_EINVAL, "ytf_operator_build: bad observation layer %d", obs_layers[l]);
    for (int k = 0; k < l; ++k)
      if (obs_layers[k] == obs_layers[l])
        return fail(c, YTF_EINVAL, "ytf_operator_build: observation layer %d given twice", obs_layers[l]);
    zoff[l] = off;
    off += n_obs[l];
    order.push_back({obs_layers[l], l});
  }
  if (off > 0 && !z_obs) return fail(c, YTF_EINVAL, "ytf_operator_build: z_obs is NULL");
  std::sort(order.begin(), order.end());
  for (auto &ol : order) {
    const double *zo = z_obs + zoff[ol.second];
    const int No = n_obs[ol.second];
    if (No == 0) continue;
    if (opt->uniform_z) {
      const double d = uniform_step(zo, No);
      if (std::isnan(d)) return fail(c, YTF_EINVAL, "build_layer_operator: z grids must be uniform");
      if (!(d >= 0)) return fail(c, YTF_EINVAL, "build_layer_operator: z_obs for layer %d must be increasing", ol.first);
    }
  }
  HIPCHK(c, hipSetDevice(c->device));

  ytf_operator *op = new ytf_operator();
  op->src = source_layer;
  op->Nx = grid->Nx;
  op->Ny = grid->Ny;
  op->Px = 2 * grid->Nx;
  op->Py = 2 * grid->Ny;
  op->dx = grid->dx;
  op->dy = grid->dy;
  op->zs.assign(z_src, z_src + n_src);
  op->flip = source_layer == 1;
  std::vector<double> zs_int = op->zs;
  if (op->flip) std::reverse(zs_int.begin(), zs_int.end());
  const int Ns = n_src;

  Build B;
  B.layers = layers;
  B.nl = n_layers;
  B.omega = omega;
  B.src = source_layer;
  B.n = n_nodes;
  B.kx = kx_dev;
  B.ky = ky_dev;
  B.W = (const cd *)w_dev;
  B.Px = op->Px;
  B.Py = op->Py;
  B.dx = grid->dx;
  B.dy = grid->dy;
  B.lateral = opt->lateral;
  B.eps = opt->nufft_eps;
  B.dz = opt->dz_box;
  B.ez_local = opt->ez_local;
  auto bail = [&](int rc) {
    hipStreamSynchronize(c->stream);
    for (auto &b : op->blocks) free_block(b);
    delete op;
    return rc;
  };
  int rc = twiddles(c, op->Px, &B.twx);
  if (!rc) rc = twiddles(c, op->Py, &B.twy);
  if (rc) return bail(rc);
  DevBuf pts, offs;
  {
    // padded offsets (i - Px/2) dx, (j - Py/2) dy; the centre offset is 0, so
    // the NUFFT ramp is 1 and the weights are W
    std::vector<double> h(op->Px + op->Py);
    for (int i = 0; i < op->Px; ++i) h[i] = (i - op->Px / 2) * grid->dx;
    for (int j = 0; j < op->Py; ++j) h[op->Px + j] = (j - op->Py / 2) * grid->dy;
    rc = dev_alloc(c, offs, sizeof(double) * h.size());
    if (!rc) rc = upload(c, offs.p, h.data(), sizeof(double) * h.size());
    if (rc) return bail(rc);
    B.xs = (double *)offs.p;
    B.ys = B.xs + op->Px;
    if (opt->lateral == YTF_LATERAL_NUFFT) {
      if (!(opt->nufft_eps >= 1e-15 && opt->nufft_eps < 1.0))
        return bail(fail(c, YTF_EINVAL, "closed_form_G: nufft_eps=%g must be in [1e-15, 1)", opt->nufft_eps));
      const size_t nb = (size_t)std::max<int64_t>(n_nodes, 1);
      rc = dev_alloc(c, pts, align256(sizeof(double) * nb) * 2 + sizeof(cd) * nb);
      if (rc) return bail(rc);
      B.xf = (double *)pts.p;
      B.yf = (double *)((char *)pts.p + align256(sizeof(double) * nb));
      B.wr = (cd *)((char *)pts.p + 2 * align256(sizeof(double) * nb));
      hipError_t e = ytf::launch_lat_points(n_nodes, kx_dev, ky_dev, B.W, grid->dx, grid->dy, 0.0, 0.0, B.xf, B.yf,
                                            B.wr, c->stream);
      if (e != hipSuccess) return bail(fail(c, YTF_EHIP, "operator build: %s", hipGetErrorString(e)));
    }
  }

  DevBuf rotb;
  if (opt->rot_n_phi > 0) {
    const int64_t nk = opt->rot_n_krho;
    const int nphi = opt->rot_n_phi;
    std::vector<double> h(2 * (size_t)nk + 2 * (size_t)nphi, 0.0);
    std::copy(opt->rot_krho, opt->rot_krho + nk, h.begin());
    for (int q = 0; q < nphi; ++q) {
      h[2 * nk + q] = std::cos(opt->rot_phi[q]);
      h[2 * nk + nphi + q] = std::sin(opt->rot_phi[q]);
    }
    rc = dev_alloc(c, rotb, sizeof(double) * h.size());
    if (!rc) rc = upload(c, rotb.p, h.data(), sizeof(double) * h.size());
    if (rc) return bail(rc);
    B.rot_nk = nk;
    B.rot_nphi = nphi;
    B.rot_kr = (const double *)rotb.p;
    B.rot_cs = B.rot_kr + 2 * nk;
  }

  Skel src_sk;
  bool have_src_sk = false;
  for (auto &ol : order) {
    const int o = ol.first, No = n_obs[ol.second];
    if (No == 0) continue;
    const auto t0 = std::chrono::steady_clock::now();
    Skel obs_sk;
    auto skel = [&](OpBlock &blk) -> int {
      if (opt->skeletons) return given_skeletons(c, opt->skeletons, source_layer, zs_int, o, blk.zo, &src_sk, &obs_sk);
      const int r = side_skeletons(c, B, o, zs_int, blk.zo, opt->skeleton_tol, &src_sk, !have_src_sk, &obs_sk);
      have_src_sk = true;
      return r;
    };
    op->blocks.emplace_back();
    OpBlock &b = op->blocks.back();
    b.obs = o;
    b.zo.assign(z_obs + zoff[ol.second], z_obs + zoff[ol.second] + No);
    if (o == source_layer) {
      const double dob = uniform_step(b.zo.data(), No);
      const bool same_step = Ns == 1 || No == 1 || (opt->uniform_z && std::fabs(std::fabs(ds) - dob) <=
                                                                           1e-9 * std::max(std::fabs(ds), dob));
      bool structured = opt->structure == YTF_STRUCTURE_STRUCTURED ||
                        (opt->structure == YTF_STRUCTURE_AUTO && same_step &&
                         structured_by_media(layers, n_layers, source_layer));
      if (structured && opt->check && opt->structure == YTF_STRUCTURE_AUTO) {
        double res = NAN;
        rc = check_structure(c, B, zs_int, b.zo, &res);
        if (rc) return bail(rc);
        b.resid = std::isnan(res) ? -1.0 : res;
        if (res > opt->check_tol) {  // NaN (grids too small to test) passes, as in the reference
          if (warnings) *warnings |= YTF_WARN_STRUCTURE_CHECK;
          structured = false;
        }
      }
      if (opt->structure == YTF_STRUCTURE_DENSE || !same_step) {
        b.kind = YTF_BLOCK_DENSE;
        rc = fill_dense(c, B, b, zs_int, YTF_TERMS_ALL);
      } else if (structured) {
        b.kind = YTF_BLOCK_STRUCTURED;
        rc = fill_toeplitz(c, B, b, zs_int, YTF_TERMS_TOEPLITZ);
        if (!rc) rc = fill_hankel(c, B, b, zs_int);
      } else {
        b.kind = YTF_BLOCK_TOEPLITZ_DENSE;
        if (skeleton) rc = skel(b);
        if (!rc) rc = fill_toeplitz(c, B, b, zs_int, YTF_TERMS_DIRECT);
        if (!rc)
          rc = fill_maybe_skeleton(c, B, b, zs_int, YTF_TERMS_SCATTERED, skeleton ? &src_sk : nullptr,
                                   skeleton ? &obs_sk : nullptr);
        if (!rc && (b.PsM || b.PoM)) b.kind = YTF_BLOCK_TOEPLITZ_SKELETON;
      }
    } else {
      b.kind = YTF_BLOCK_DENSE;
      if (skeleton) rc = skel(b);
      if (!rc)
        rc = fill_maybe_skeleton(c, B, b, zs_int, YTF_TERMS_ALL, skeleton ? &src_sk : nullptr,
                                 skeleton ? &obs_sk : nullptr);
      if (!rc && (b.PsM || b.PoM)) b.kind = YTF_BLOCK_SKELETON;
    }
    if (!rc) rc = merge_planes(c, b, Ns, (int64_t)op->Px * op->Py);
    if (rc) return bail(rc);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    b.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = op;
  return YTF_OK;
}

int ytf_operator_destroy(ytf_ctx *c, ytf_operator *op) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_operator_destroy: ctx is NULL");
  if (!op) return YTF_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &b : op->blocks) free_block(b);
  if (op->work) hipFree(op->work);
  if (op->stage) hipFree(op->stage);
  delete op;
  return YTF_OK;
}

int ytf_operator_num_blocks(const ytf_operator *op) { return op ? (int)op->blocks.size() : 0; }

int ytf_operator_block_info(ytf_ctx *c, const ytf_operator *op, int block, ytf_block_info *info) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_operator_block_info: ctx is NULL");
  if (!op || !info || block < 0 || block >= (int)op->blocks.size())
    return fail(c, YTF_EINVAL, "ytf_operator_block_info: bad block %d", block);
  const OpBlock &b = op->blocks[block];
  info->obs_layer = b.obs;
  info->kind = b.kind;
  info->n_obs = (int)b.zo.size();
  info->n_src_skel = b.D ? b.nDs : (int)op->zs.size();
  info->n_obs_skel = b.D ? b.nDo : (int)b.zo.size();
  info->n_toeplitz = b.nT;
  info->n_hankel = b.nH;
  info->n_dense = b.D ? b.nDs * b.nDo : 0;
  info->bytes = b.bytes;
  info->build_ms = b.build_ms;
  info->check_residual = b.resid;
  return YTF_OK;
}

int ytf_operator_skeleton(ytf_ctx *c, const ytf_operator *op, int block, int side, int32_t *n_planes,
                          int32_t *idx, double *P, double *err) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_operator_skeleton: ctx is NULL");
  if (!op || block < 0 || block >= (int)op->blocks.size() || (side != 0 && side != 1) || !n_planes)
    return fail(c, YTF_EINVAL, "ytf_operator_skeleton: bad block %d or side %d", block, side);
  const OpBlock &b = op->blocks[block];
  const int N = side == 0 ? (int)op->zs.size() : (int)b.zo.size();
  const std::vector<int> &ix = side == 0 ? b.sidx : b.oidx;
  const std::vector<zc> &Pm = side == 0 ? b.Ps : b.Po;
  const int k = ix.empty() ? N : (int)ix.size();
  *n_planes = k;
  if (idx)
    for (int q = 0; q < k; ++q) idx[q] = ix.empty() ? q : ix[q];
  if (P) {
    for (int e = 0; e < N * k; ++e) {
      const zc v = Pm.empty() ? zc((e % N) == (e / N) ? 1.0 : 0.0, 0.0) : Pm[e];
      P[2 * e] = v.real();
      P[2 * e + 1] = v.imag();
    }
  }
  if (err) *err = side == 0 ? b.serr : b.oerr;
  return YTF_OK;
}

int ytf_operator_ghat(ytf_ctx *c, const ytf_operator *op, int block, int part, int slot, double *ghat_dev) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_operator_ghat: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!op || !ghat_dev || block < 0 || block >= (int)op->blocks.size())
    return fail(c, YTF_EINVAL, "ytf_operator_ghat: bad block %d", block);
  const OpBlock &b = op->blocks[block];
  const cd *base = part == 0 ? b.T : part == 1 ? b.H : part == 2 ? b.D : nullptr;
  const int ns = part == 0 ? b.nT : part == 1 ? b.nH : b.nDs * b.nDo;
  if (!base || slot < 0 || slot >= ns) return fail(c, YTF_EINVAL, "ytf_operator_ghat: no plane %d of part %d", slot, part);
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t nn = (int64_t)op->Px * op->Py;
  hipError_t e = ytf::launch_op_export(op->Px, op->Py, base + (int64_t)slot * 9 * nn, (cd *)ghat_dev, c->stream);
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_operator_ghat: %s", hipGetErrorString(e));
  return YTF_OK;
}

}  // extern "C"

namespace {

int apply_locked(ytf_ctx *c, ytf_operator *op, const double *P_dev, double *const *E_dev) {
  if (!op || !P_dev || (!op->blocks.empty() && !E_dev)) return fail(c, YTF_EINVAL, "ytf_operator_apply: NULL argument");
  for (size_t b = 0; b < op->blocks.size(); ++b)
    if (!E_dev[b]) return fail(c, YTF_EINVAL, "ytf_operator_apply: E for block %zu is NULL", b);
  HIPCHK(c, hipSetDevice(c->device));
  const int Ns = (int)op->zs.size(), Nx = op->Nx, Ny = op->Ny, Px = op->Px, Py = op->Py;
  const int64_t nn = (int64_t)Px * Py, nxy = (int64_t)Nx * Ny;
  int no_max = 0, rs = 0, ro_max = 0;
  const OpBlock *fold = nullptr;  // every skeleton block shares the source skeleton
  for (auto &b : op->blocks) {
    no_max = std::max(no_max, (int)b.zo.size());
    if (b.PsM && !fold) {
      fold = &b;
      rs = b.nDs;
    }
    if (b.PoM) ro_max = std::max(ro_max, b.nDo);
  }
  const size_t need = sizeof(cd) * 3 * (size_t)nn * (size_t)(Ns + no_max + rs + ro_max);
  if (need > op->work_bytes) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (op->work) hipFree(op->work);
    op->work = nullptr;
    op->work_bytes = 0;
    HIPCHK(c, hipMalloc(&op->work, need));
    op->work_bytes = need;
  }
  cd *J = (cd *)op->work, *Ew = J + 3 * (size_t)nn * Ns, *Js = Ew + 3 * (size_t)nn * no_max,
     *Es = Js + 3 * (size_t)nn * rs;
  cd *twx = nullptr, *twy = nullptr;
  int rc = twiddles(c, Px, &twx);
  if (!rc) rc = twiddles(c, Py, &twy);
  if (rc) return rc;
  // source spectra J[k][c] of the zero-padded moments, k in increasing
  // cavity-frame depth (source layer 1 reversed)
  const cd *P = (const cd *)P_dev;
  const cd *pbase = op->flip ? P + (int64_t)(Ns - 1) * nxy : P;
  const int64_t pa = op->flip ? -nxy : nxy;
  bool fused = !op->blocks.empty();
  for (auto &b : op->blocks) fused = fused && single_plane(b, Ns);
  fused = fused && ytf::opy_mix1_supported(Py, Px);
  if (fused) {
    // one plane per block: FFT along x, then per block the fused y FFT,
    // 3x3 product and inverse y FFT, and the inverse x FFT
    hipError_t e = timed(c, "op_fft_fwd", [&] {
      return ytf::launch_opx(false, Px, pbase, 3, pa, (int64_t)Ns * nxy, Nx, Nx, Ny, J, Py, nn, 3, twx, c->stream);
    });
    for (size_t bi = 0; bi < op->blocks.size() && e == hipSuccess; ++bi) {
      const OpBlock &b = op->blocks[bi];
      e = timed(c, "op_mix", [&] {
        return ytf::launch_opy_mix1(Py, J, Ew, Px, b.T ? b.T : b.D, Ny, Ny, twy, c->stream);
      });
      if (e == hipSuccess)
        e = timed(c, "op_fft_inv", [&] {
          return ytf::launch_opx(true, Px, Ew, 3, nxy, nxy, Nx, Nx, Ny, (cd *)E_dev[bi], Py, nn, 3, twx, c->stream);
        });
    }
    if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_operator_apply: launch failed: %s", hipGetErrorString(e));
    return YTF_OK;
  }
  hipError_t e = timed(c, "op_fft_fwd", [&] {
    hipError_t r = ytf::launch_opx(false, Px, pbase, 3, pa, (int64_t)Ns * nxy, Nx, Nx, Ny, J, Py, nn, 3 * Ns, twx,
                                   c->stream);
    if (r == hipSuccess) r = ytf::launch_opy(false, Py, J, Px, nn, 3 * Ns, Ny, Py, 1.0, 0, twy, c->stream);
    return r;
  });
  if (e == hipSuccess && fold)  // source planes folded onto the skeleton planes (_zfold)
    e = timed(c, "op_mix", [&] { return ytf::launch_op_zcomb(nn, J, Ns, fold->PsM, Js, rs, 0, c->stream); });
  for (size_t bi = 0; bi < op->blocks.size() && e == hipSuccess; ++bi) {
    const OpBlock &b = op->blocks[bi];
    const int No = (int)b.zo.size();
    e = timed(c, "op_mix", [&] {
      hipError_t r = hipSuccess;
      bool acc = false;
      if (b.T) {  // Toeplitz planes, with the Hankel planes in the same pass
        r = ytf::launch_op_mix(nn, J, Ns, Ew, No, b.T, b.H, ytf::OP_MIX_TOEPLITZ, acc, c->stream);
        acc = true;
      } else if (b.H) {
        r = ytf::launch_op_mix(nn, J, Ns, Ew, No, b.H, nullptr, ytf::OP_MIX_HANKEL, acc, c->stream);
        acc = true;
      }
      if (r == hipSuccess && b.D) {
        const cd *Jin = b.PsM ? Js : J;
        if (b.PoM) {  // skeleton observation planes, interpolated to all planes (_zexpand_add!)
          r = ytf::launch_op_mix(nn, Jin, b.nDs, Es, b.nDo, b.D, nullptr, ytf::OP_MIX_DENSE, 0, c->stream);
          if (r == hipSuccess) r = ytf::launch_op_zcomb(nn, Es, b.nDo, b.PoM, Ew, No, acc, c->stream);
        } else {
          r = ytf::launch_op_mix(nn, Jin, b.nDs, Ew, No, b.D, nullptr, ytf::OP_MIX_DENSE, acc, c->stream);
        }
        acc = true;
      }
      return r;
    });
    if (e != hipSuccess) break;
    double *Eb = E_dev[bi];
    e = timed(c, "op_fft_inv", [&] {
      hipError_t r = ytf::launch_opy(true, Py, Ew, Px, nn, 3 * No, Py, Ny, 1.0, 0, twy, c->stream);
      if (r == hipSuccess)
        r = ytf::launch_opx(true, Px, Ew, 3, nxy, (int64_t)No * nxy, Nx, Nx, Ny, (cd *)Eb, Py, nn, 3 * No, twx,
                            c->stream);
      return r;
    });
  }
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_operator_apply: launch failed: %s", hipGetErrorString(e));
  return YTF_OK;
}

const char *block_kind_name(int k) {
  static const char *names[] = {":structured", ":toeplitz_dense", ":dense", ":skeleton", ":toeplitz_skeleton"};
  return (k >= 0 && k < 5) ? names[k] : ":unknown";
}

}  // namespace

extern "C" {

int ytf_operator_apply(ytf_ctx *c, ytf_operator *op, const double *P_dev, double *const *E_dev) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_operator_apply: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  int rc = YTF_OK;
  timed(c, "op_apply", [&] {
    rc = apply_locked(c, op, P_dev, E_dev);
    return hipSuccess;
  });
  return rc;
}

int ytf_operator_apply_host(ytf_ctx *c, ytf_operator *op, const double *P_host, double *const *E_host) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_operator_apply_host: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!op || !P_host || (!op->blocks.empty() && !E_host))
    return fail(c, YTF_EINVAL, "ytf_operator_apply_host: NULL argument");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t nxy3 = 3 * (size_t)op->Nx * op->Ny;
  const size_t pb = align256(sizeof(cd) * nxy3 * op->zs.size());
  std::vector<size_t> eoff;
  size_t need = pb;
  for (auto &b : op->blocks) {
    eoff.push_back(need);
    need += align256(sizeof(cd) * nxy3 * b.zo.size());
  }
  if (need > op->stage_bytes) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (op->stage) hipFree(op->stage);
    op->stage = nullptr;
    op->stage_bytes = 0;
    HIPCHK(c, hipMalloc(&op->stage, need));
    op->stage_bytes = need;
  }
  char *base = (char *)op->stage;
  HIPCHK(c, hipMemcpyAsync(base, P_host, sizeof(cd) * nxy3 * op->zs.size(), hipMemcpyHostToDevice, c->stream));
  std::vector<double *> E(op->blocks.size());
  for (size_t b = 0; b < E.size(); ++b) E[b] = (double *)(base + eoff[b]);
  int rc = apply_locked(c, op, (const double *)base, E.data());
  if (rc) return rc;
  for (size_t b = 0; b < E.size(); ++b) {
    if (!E_host[b]) return fail(c, YTF_EINVAL, "ytf_operator_apply_host: E for block %zu is NULL", b);
    HIPCHK(c, hipMemcpyAsync(E_host[b], E[b], sizeof(cd) * nxy3 * op->blocks[b].zo.size(), hipMemcpyDeviceToHost,
                             c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return YTF_OK;
}

int ytf_operator_accumulate(ytf_ctx *c, ytf_operator *dst, const ytf_operator *src) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_operator_accumulate: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!dst || !src) return fail(c, YTF_EINVAL, "ytf_operator_accumulate: NULL argument");
  if (dst == src) return fail(c, YTF_EINVAL, "ytf_operator_accumulate: dst and src are the same operator");
  if (dst->src != src->src || dst->Nx != src->Nx || dst->Ny != src->Ny || dst->dx != src->dx ||
      dst->dy != src->dy || !same_planes(dst->zs, src->zs.data(), (int)src->zs.size()) ||
      dst->blocks.size() != src->blocks.size())
    return fail(c, YTF_EINVAL, "build_layer_operator_grouped: groups have different grids, source planes or blocks");
  for (size_t i = 0; i < dst->blocks.size(); ++i) {
    const OpBlock &a = dst->blocks[i], &b = src->blocks[i];
    if (a.obs != b.obs || !same_planes(a.zo, b.zo.data(), (int)b.zo.size()))
      return fail(c, YTF_EINVAL, "build_layer_operator_grouped: groups have different observation planes");
    if (a.kind != b.kind)
      return fail(c, YTF_EINVAL, "build_layer_operator_grouped: groups chose different block kinds (%s vs %s)",
                  block_kind_name(a.kind), block_kind_name(b.kind));
    if (a.sidx != b.sidx || a.oidx != b.oidx || a.Ps != b.Ps || a.Po != b.Po)
      return fail(c, YTF_EINVAL, "build_layer_operator_grouped: groups have different z skeletons");
    if (a.nT != b.nT || a.nH != b.nH || a.nDs != b.nDs || a.nDo != b.nDo || !a.T != !b.T || !a.H != !b.H ||
        !a.D != !b.D)
      return fail(c, YTF_EINVAL, "build_layer_operator_grouped: groups stored different planes");
  }
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t nn9 = 9 * (int64_t)dst->Px * dst->Py;
  for (size_t i = 0; i < dst->blocks.size(); ++i) {
    OpBlock &a = dst->blocks[i];
    const OpBlock &b = src->blocks[i];
    int rc = add_planes(c, a.T, b.T, a.nT * nn9);
    if (!rc) rc = add_planes(c, a.H, b.H, a.nH * nn9);
    if (!rc) rc = add_planes(c, a.D, b.D, (int64_t)a.nDs * a.nDo * nn9);
    if (rc) return rc;
    a.build_ms += b.build_ms;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return YTF_OK;
}

int ytf_z_skeletons(ytf_ctx *c, const ytf_layer *layers, int n_layers, double omega, int source_layer,
                    int64_t n_nodes, const double *kx_dev, const double *ky_dev, const double *z_src, int n_src,
                    int n_obs_layers, const int32_t *obs_layers, const int32_t *n_obs, const double *z_obs,
                    double tol, ytf_skeletons **out) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_z_skeletons: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!out || !layers) return fail(c, YTF_EINVAL, "ytf_z_skeletons: NULL argument");
  *out = nullptr;
  if (source_layer < 1 || source_layer > n_layers - 1)
    return fail(c, YTF_EINVAL, "ytf_z_skeletons: source_layer=%d must be in 1:%d", source_layer, n_layers - 1);
  if (n_src < 1 || !z_src) return fail(c, YTF_EINVAL, "z_skeletons: need at least one source plane");
  if (n_nodes < 0 || (n_nodes > 0 && (!kx_dev || !ky_dev))) return fail(c, YTF_EINVAL, "ytf_z_skeletons: bad nodes");
  if (!(tol > 0)) return fail(c, YTF_EINVAL, "z_skeletons: tol must be > 0");
  if (n_obs_layers < 0 || (n_obs_layers > 0 && (!obs_layers || !n_obs)))
    return fail(c, YTF_EINVAL, "ytf_z_skeletons: bad observation layers");
  std::vector<std::pair<int, int>> order;
  std::vector<int64_t> zoff(n_obs_layers);
  int64_t off = 0;
  for (int l = 0; l < n_obs_layers; ++l) {
    if (obs_layers[l] < 1 || obs_layers[l] > n_layers || n_obs[l] < 0)
      return fail(c, YTF_EINVAL, "ytf_z_skeletons: bad observation layer %d", obs_layers[l]);
    for (int k = 0; k < l; ++k)
      if (obs_layers[k] == obs_layers[l])
        return fail(c, YTF_EINVAL, "ytf_z_skeletons: observation layer %d given twice", obs_layers[l]);
    zoff[l] = off;
    off += n_obs[l];
    order.push_back({obs_layers[l], l});
  }
  if (off > 0 && !z_obs) return fail(c, YTF_EINVAL, "ytf_z_skeletons: z_obs is NULL");
  std::sort(order.begin(), order.end());
  HIPCHK(c, hipSetDevice(c->device));
  Build B;
  B.layers = layers;
  B.nl = n_layers;
  B.omega = omega;
  B.src = source_layer;
  B.n = n_nodes;
  B.kx = kx_dev;
  B.ky = ky_dev;
  ytf_skeletons *sk = new ytf_skeletons();
  sk->src = source_layer;
  sk->zs_int.assign(z_src, z_src + n_src);
  if (source_layer == 1) std::reverse(sk->zs_int.begin(), sk->zs_int.end());
  bool have_src = false;
  for (auto &ol : order) {
    std::vector<double> zo(z_obs + zoff[ol.second], z_obs + zoff[ol.second] + n_obs[ol.second]);
    Skel os;
    int rc = side_skeletons(c, B, ol.first, sk->zs_int, zo, tol, &sk->src_sk, !have_src, &os);
    if (rc) {
      delete sk;
      return rc;
    }
    have_src = true;
    sk->obs.push_back(ol.first);
    sk->zo.push_back(zo);
    sk->obs_sk.push_back(os);
  }
  if (!have_src) {  // no observation layer: the source side alone
    Skel os;
    int rc = side_skeletons(c, B, source_layer, sk->zs_int, std::vector<double>(), tol, &sk->src_sk, true, &os);
    if (rc) {
      delete sk;
      return rc;
    }
  }
  *out = sk;
  return YTF_OK;
}

int ytf_z_skeletons_q(ytf_ctx *c, const ytf_layer *layers, int n_layers, int source_layer, int64_t n_nodes,
                      const double *Qsrc_dev, const double *zs_int, int n_src, int n_obs_layers,
                      const int32_t *obs_layers, const double *const *Qobs_dev, const int32_t *n_obs,
                      const double *z_obs, double tol, ytf_skeletons **out) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_z_skeletons_q: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!out || !layers) return fail(c, YTF_EINVAL, "ytf_z_skeletons_q: NULL argument");
  *out = nullptr;
  if (source_layer < 1 || source_layer > n_layers - 1)
    return fail(c, YTF_EINVAL, "ytf_z_skeletons_q: source_layer=%d must be in 1:%d", source_layer, n_layers - 1);
  if (n_src < 1 || !zs_int) return fail(c, YTF_EINVAL, "z_skeletons: need at least one source plane");
  if (n_nodes < 0 || (n_nodes > 0 && !Qsrc_dev)) return fail(c, YTF_EINVAL, "ytf_z_skeletons_q: bad Qsrc");
  if (!(tol > 0)) return fail(c, YTF_EINVAL, "z_skeletons: tol must be > 0");
  if (n_obs_layers < 0 || (n_obs_layers > 0 && (!obs_layers || !n_obs || !Qobs_dev)))
    return fail(c, YTF_EINVAL, "ytf_z_skeletons_q: bad observation layers");
  std::vector<std::pair<int, int>> order;
  std::vector<int64_t> zoff(n_obs_layers);
  int64_t off = 0;
  for (int l = 0; l < n_obs_layers; ++l) {
    if (obs_layers[l] < 1 || obs_layers[l] > n_layers || n_obs[l] < 0 || (n_nodes > 0 && !Qobs_dev[l]))
      return fail(c, YTF_EINVAL, "ytf_z_skeletons_q: bad observation layer %d", obs_layers[l]);
    for (int k = 0; k < l; ++k)
      if (obs_layers[k] == obs_layers[l])
        return fail(c, YTF_EINVAL, "ytf_z_skeletons_q: observation layer %d given twice", obs_layers[l]);
    zoff[l] = off;
    off += n_obs[l];
    order.push_back({obs_layers[l], l});
  }
  if (off > 0 && !z_obs) return fail(c, YTF_EINVAL, "ytf_z_skeletons_q: z_obs is NULL");
  std::sort(order.begin(), order.end());
  HIPCHK(c, hipSetDevice(c->device));
  ytf_skeletons *sk = new ytf_skeletons();
  sk->src = source_layer;
  sk->zs_int.assign(zs_int, zs_int + n_src);
  int rc = src_skeleton(c, layers, source_layer, n_nodes, (const cd *)Qsrc_dev, sk->zs_int, tol, &sk->src_sk);
  for (size_t i = 0; i < order.size() && !rc; ++i) {
    const int l = order[i].second;
    std::vector<double> zo(z_obs + zoff[l], z_obs + zoff[l] + n_obs[l]);
    Skel os;
    rc = obs_skeleton(c, layers, n_layers, order[i].first, n_nodes, (const cd *)Qobs_dev[l], zo, tol, &os);
    sk->obs.push_back(order[i].first);
    sk->zo.push_back(zo);
    sk->obs_sk.push_back(os);
  }
  if (rc) {
    delete sk;
    return rc;
  }
  *out = sk;
  return YTF_OK;
}

int ytf_skeletons_get(ytf_ctx *c, const ytf_skeletons *sk, int layer, int32_t *n_planes, int32_t *n_skel,
                      int32_t *idx, double *P, double *err) {
  if (!sk) return fail(c, YTF_EINVAL, "ytf_skeletons_get: NULL skeleton set");
  const Skel *s = nullptr;
  if (layer == 0) {
    s = &sk->src_sk;
  } else {
    for (size_t l = 0; l < sk->obs.size(); ++l)
      if (sk->obs[l] == layer) s = &sk->obs_sk[l];
  }
  if (!s) return fail(c, YTF_EINVAL, "ytf_skeletons_get: no side for layer %d", layer);
  const int N = s->N, k = (int)s->idx.size();
  if (n_planes) *n_planes = N;
  if (n_skel) *n_skel = k;
  if (idx)
    for (int q = 0; q < k; ++q) idx[q] = s->idx[q];
  if (P) std::memcpy(P, s->P.data(), sizeof(zc) * (size_t)N * k);
  if (err) *err = s->err;
  return YTF_OK;
}

int ytf_skeletons_destroy(ytf_ctx *c, ytf_skeletons *sk) {
  (void)c;
  delete sk;
  return YTF_OK;
}

int ytf_buffer_alloc(ytf_ctx *c, size_t bytes, void **dev) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_buffer_alloc: ctx is NULL");
  if (!dev) return fail(c, YTF_EINVAL, "ytf_buffer_alloc: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMalloc(dev, bytes < 256 ? 256 : bytes));
  return YTF_OK;
}

int ytf_buffer_free(ytf_ctx *c, void *dev) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_buffer_free: ctx is NULL");
  if (!dev) return YTF_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipFree(dev));
  return YTF_OK;
}

int ytf_buffer_upload(ytf_ctx *c, void *dev, const void *host, size_t bytes) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_buffer_upload: ctx is NULL");
  if (bytes && (!dev || !host)) return fail(c, YTF_EINVAL, "ytf_buffer_upload: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(c, hipSetDevice(c->device));
  if (bytes) HIPCHK(c, hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return YTF_OK;
}

int ytf_buffer_download(ytf_ctx *c, void *host, const void *dev, size_t bytes) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_buffer_download: ctx is NULL");
  if (bytes && (!dev || !host)) return fail(c, YTF_EINVAL, "ytf_buffer_download: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(c, hipSetDevice(c->device));
  if (bytes) HIPCHK(c, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return YTF_OK;
}

}  // extern "C"
