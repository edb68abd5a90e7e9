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

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

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
  std::vector<int> sidx, oidx;
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
  // device: NUFFT points and weights, padded offsets
  double *xf = nullptr, *yf = nullptr;
  cd *wr = nullptr;
  double *xs = nullptr, *ys = nullptr;
  cd *twx = nullptr, *twy = nullptr;
};

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
    rc = green_nodes_locked(c, B.layers, B.nl, B.omega, B.kx, B.ky, B.n, B.src, obs_layer, np, zs.data(), zo.data(),
                            (double *)g, nullptr, terms);
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
  rc = green_nodes_locked(c, B.layers, B.nl, B.omega, B.kx, B.ky, B.n, B.src, B.src, 4, zsv, zov, (double *)gb.p,
                          nullptr, YTF_TERMS_SCATTERED);
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

int fill_dense(ytf_ctx *c, const Build &B, OpBlock &b, const std::vector<double> &zs, int terms) {
  const int64_t nn = (int64_t)B.Px * B.Py;
  b.nDs = (int)zs.size();
  b.nDo = (int)b.zo.size();
  int rc = alloc_planes(c, &b.D, b.nDs * b.nDo, nn, &b.bytes);
  if (rc) return rc;
  std::vector<Req> reqs;
  for (int j = 0; j < b.nDo; ++j)
    for (int i = 0; i < b.nDs; ++i) reqs.push_back({zs[i], b.zo[j], b.D + (int64_t)(i + b.nDs * j) * 9 * nn});
  return fill_planes(c, B, b.obs, reqs, terms);
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
  if (b.D && b.T && Ns == 1 && No == 1 && b.nT == 1) {
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
  return Ns == 1 && b.zo.size() == 1 && !b.H && ((b.T && !b.D && b.nT == 1) || (b.D && !b.T));
}

}  // namespace

extern "C" {

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
  if (!pow2(grid->Nx) || !pow2(grid->Ny) || grid->Nx < 4 || grid->Ny < 4 || grid->Nx > 1024 || grid->Ny > 1024)
    return fail(c, YTF_EINVAL, "ytf_operator_build: grid Nx=%d, Ny=%d must be powers of two in 4..1024", grid->Nx,
                grid->Ny);
  if (!(grid->dx > 0) || !(grid->dy > 0)) return fail(c, YTF_EINVAL, "ytf_operator_build: grid pitch must be > 0");
  if (n_src < 1 || !z_src) return fail(c, YTF_EINVAL, "build_layer_operator: need at least one source plane");
  if (n_obs_layers < 0 || (n_obs_layers > 0 && (!obs_layers || !n_obs)))
    return fail(c, YTF_EINVAL, "ytf_operator_build: bad observation layers");
  if (n_nodes < 0 || (n_nodes > 0 && (!kx_dev || !ky_dev || !w_dev)))
    return fail(c, YTF_EINVAL, "ytf_operator_build: bad nodes");
  if (source_layer < 1 || source_layer > n_layers - 1)
    return fail(c, YTF_EINVAL, "ytf_operator_build: source_layer=%d must be in 1:%d", source_layer, n_layers - 1);
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

  for (auto &ol : order) {
    const int o = ol.first, No = n_obs[ol.second];
    if (No == 0) continue;
    const auto t0 = std::chrono::steady_clock::now();
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
        if (skeleton) return bail(fail(c, YTF_EINVAL, "ytf_operator_build: skeleton planes are not available"));
        b.kind = YTF_BLOCK_TOEPLITZ_DENSE;
        rc = fill_toeplitz(c, B, b, zs_int, YTF_TERMS_DIRECT);
        if (!rc) rc = fill_dense(c, B, b, zs_int, YTF_TERMS_SCATTERED);
      }
    } else {
      if (skeleton) return bail(fail(c, YTF_EINVAL, "ytf_operator_build: skeleton planes are not available"));
      b.kind = YTF_BLOCK_DENSE;
      rc = fill_dense(c, B, b, zs_int, YTF_TERMS_ALL);
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

int ytf_operator_apply(ytf_ctx *c, ytf_operator *op, const double *P_dev, double *const *E_dev) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_operator_apply: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!op || !P_dev || (!op->blocks.empty() && !E_dev)) return fail(c, YTF_EINVAL, "ytf_operator_apply: NULL argument");
  for (size_t b = 0; b < op->blocks.size(); ++b)
    if (!E_dev[b]) return fail(c, YTF_EINVAL, "ytf_operator_apply: E for block %zu is NULL", b);
  HIPCHK(c, hipSetDevice(c->device));
  const int Ns = (int)op->zs.size(), Nx = op->Nx, Ny = op->Ny, Px = op->Px, Py = op->Py;
  const int64_t nn = (int64_t)Px * Py, nxy = (int64_t)Nx * Ny;
  int no_max = 0;
  for (auto &b : op->blocks) no_max = std::max(no_max, (int)b.zo.size());
  const size_t need = sizeof(cd) * 3 * (size_t)nn * (size_t)(Ns + no_max);
  if (need > op->work_bytes) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (op->work) hipFree(op->work);
    op->work = nullptr;
    op->work_bytes = 0;
    HIPCHK(c, hipMalloc(&op->work, need));
    op->work_bytes = need;
  }
  cd *J = (cd *)op->work, *Ew = J + 3 * (size_t)nn * Ns;
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
  for (size_t bi = 0; bi < op->blocks.size() && e == hipSuccess; ++bi) {
    const OpBlock &b = op->blocks[bi];
    const int No = (int)b.zo.size();
    e = timed(c, "op_mix", [&] {
      hipError_t r = hipSuccess;
      bool acc = false;
      if (b.T) {
        r = ytf::launch_op_mix(nn, J, Ns, Ew, No, b.T, ytf::OP_MIX_TOEPLITZ, acc, c->stream);
        acc = true;
      }
      if (r == hipSuccess && b.H) {
        r = ytf::launch_op_mix(nn, J, Ns, Ew, No, b.H, ytf::OP_MIX_HANKEL, acc, c->stream);
        acc = true;
      }
      if (r == hipSuccess && b.D) {
        r = ytf::launch_op_mix(nn, J, Ns, Ew, No, b.D, ytf::OP_MIX_DENSE, acc, c->stream);
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

}  // extern "C"
