// capi.hip -- implementation of the C-ABI declared in include/ytf.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <map>
#include <tuple>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"

static_assert(sizeof(LayerDev) == sizeof(ytf_layer), "ytf_layer and LayerDev must share one layout");

extern "C" {

const char *ytf_version(void) { return "yellowthinfilms_amd 0.1 (gfx950)"; }

int ytf_ctx_create(int device, ytf_ctx **out) {
  if (!out) return fail(nullptr, YTF_EINVAL, "ytf_ctx_create: out is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(nullptr, YTF_EHIP, "ytf_ctx_create: no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= n) return fail(nullptr, YTF_EINVAL, "ytf_ctx_create: device %d out of range 0..%d", device, n - 1);
  ytf_ctx *c = new ytf_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(nullptr, YTF_EHIP, "ytf_ctx_create: stream creation failed");
  }
  c->stream = c->own;
  if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) c->n_cu = 0;
  std::string err;
  if (int rc = noalloc_init(device, &err)) {
    hipStreamDestroy(c->own);
    delete c;
    return fail(nullptr, rc, "ytf_ctx_create: %s", err.c_str());
  }
  *out = c;
  return YTF_OK;
}

int ytf_ctx_destroy(ytf_ctx *c) {
  if (!c) return YTF_OK;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  for (auto &kv : c->tw) hipFree(kv.second);
  for (auto &kv : c->nu_deconv) hipFree(kv.second);
  for (auto &kv : c->nu_horner) hipFree(kv.second);
  for (auto &m : c->marks) {
    hipEventDestroy(m.a);
    hipEventDestroy(m.b);
  }
  for (auto e : c->pool) hipEventDestroy(e);
  for (auto &s : c->small_slots)
    if (s.dev) hipFree(s.dev);
  if (c->work) hipFree(c->work);
  if (c->aux) hipFree(c->aux);
  if (c->redo) hipFree(c->redo);
  if (c->pinned) hipHostFree(c->pinned);
  if (c->own) hipStreamDestroy(c->own);
  delete c;
  return YTF_OK;
}

const char *ytf_last_error(const ytf_ctx *c) { return c ? c->err.c_str() : g_err.c_str(); }

int ytf_ctx_set_stream(ytf_ctx *c, void *s) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_ctx_set_stream: ctx is NULL");
  // NULL is the HIP null (legacy default) stream, the one torch's default
  // current stream wraps; work is then ordered with everything on it
  c->stream = (hipStream_t)s;
  return YTF_OK;
}

int ytf_ctx_synchronize(ytf_ctx *c) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_ctx_synchronize: ctx is NULL");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return YTF_OK;
}

int ytf_green_nodes(ytf_ctx *c, const ytf_layer *layers, int n_layers, double omega, const double *kx_dev,
                    const double *ky_dev, int64_t n_nodes, int source_layer, int obs_layer, int n_pairs,
                    const double *zsrc, const double *zobs, double *ghat_dev, const ytf_node_cache *cache,
                    int terms) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_green_nodes: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  return green_nodes_locked(c, layers, n_layers, omega, kx_dev, ky_dev, n_nodes, source_layer, obs_layer, n_pairs,
                            zsrc, zobs, ghat_dev, cache, terms, 0.0, 1);
}

int ytf_green_nodes_box(ytf_ctx *c, const ytf_layer *layers, int n_layers, double omega, const double *kx_dev,
                        const double *ky_dev, int64_t n_nodes, int source_layer, int obs_layer, int n_pairs,
                        const double *zsrc, const double *zobs, double *ghat_dev, const ytf_node_cache *cache,
                        int terms, double dz_box, int ez_local) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_green_nodes_box: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  return green_nodes_locked(c, layers, n_layers, omega, kx_dev, ky_dev, n_nodes, source_layer, obs_layer, n_pairs,
                            zsrc, zobs, ghat_dev, cache, terms, dz_box, ez_local);
}

}  // extern "C"

// Closed-form layers (isotropic/uniaxial eps, isotropic mu) and layers
// without xz/yz coupling never reach the general eigen-solver.
static bool stack_needs_general(const ytf_layer *layers, int n_layers) {
  for (int l = 0; l < n_layers; ++l) {
    const ytf_layer &Ly = layers[l];
    if ((Ly.eps_kind == 0 || Ly.eps_kind == 1) && Ly.mu_kind == 0) continue;
    const int idx[4] = {2, 5, 6, 7};
    for (int t = 0; t < 4; ++t)
      if (Ly.eps[2 * idx[t]] != 0.0 || Ly.eps[2 * idx[t] + 1] != 0.0 || Ly.mu[2 * idx[t]] != 0.0 ||
          Ly.mu[2 * idx[t] + 1] != 0.0)
        return true;
  }
  return false;
}

// Validate the arguments of a node solve, send the stack and depth pairs to
// the device (when they differ from what its copy holds) and fill *P.  With
// n_nodes == 0 nothing is sent and P->nn = 0.
static int node_params_locked(ytf_ctx *c, const ytf_layer *layers, int n_layers, double omega, const double *kx_dev,
                              const double *ky_dev, int64_t n_nodes, int source_layer, int obs_layer, int n_pairs,
                              const double *zsrc, const double *zobs, double *ghat_dev, const ytf_node_cache *cache,
                              int terms, double dz_box, int ez_local, ytf::NodeParams *P_out) {
  P_out->nn = 0;
  if (!layers || n_layers < 2) return fail(c, YTF_EINVAL, "ytf_green_nodes: need at least 2 layers (got %d)", n_layers);
  if (source_layer < 1 || source_layer > n_layers - 1)
    return fail(c, YTF_EINVAL,
                "ytf_green_nodes: source_layer=%d must be in 1:%d (the last, semi-infinite layer cannot host a source)",
                source_layer, n_layers - 1);
  if (obs_layer < 1 || obs_layer > n_layers)
    return fail(c, YTF_EINVAL, "ytf_green_nodes: obs_layer=%d out of range 1:%d", obs_layer, n_layers);
  if (n_nodes < 0 || n_nodes > INT32_MAX) return fail(c, YTF_EINVAL, "ytf_green_nodes: n_nodes must be in 0:2^31-1");
  if (terms < YTF_TERMS_ALL || terms > YTF_TERMS_HANKEL)
    return fail(c, YTF_EINVAL, "closed_form_G: terms must be :all, :direct, :scattered, :toeplitz or :hankel (got %d)",
                terms);
  if ((terms == YTF_TERMS_DIRECT || terms == YTF_TERMS_TOEPLITZ || terms == YTF_TERMS_HANKEL) &&
      obs_layer != source_layer)
    return fail(c, YTF_EINVAL, "closed_form_G: terms=:%s needs obs_layer == source_layer (%d), got %d",
                terms == YTF_TERMS_DIRECT ? "direct" : terms == YTF_TERMS_TOEPLITZ ? "toeplitz" : "hankel",
                source_layer, obs_layer);
  if (n_nodes > 0 && (!kx_dev || !ky_dev)) return fail(c, YTF_EINVAL, "ytf_green_nodes: kx/ky are NULL");
  if (n_pairs < 0 || (n_pairs > 0 && (!zsrc || !zobs))) return fail(c, YTF_EINVAL, "ytf_green_nodes: bad depth pairs");
  if (ghat_dev && n_pairs == 0) return fail(c, YTF_EINVAL, "ytf_green_nodes: ghat requested with no depth pairs");
  if (!(omega > 0) || !std::isfinite(omega)) return fail(c, YTF_EINVAL, "ytf_green_nodes: omega must be > 0");
  for (int l = 0; l < n_layers; ++l) {
    if (layers[l].eps_kind < 0 || layers[l].eps_kind > 3 || layers[l].mu_kind < 0 || layers[l].mu_kind > 1)
      return fail(c, YTF_EINVAL, "ytf_green_nodes: layer %d has an unknown material kind", l + 1);
  }
  const double ds = layers[source_layer - 1].d;
  if (!(dz_box >= 0.0) || !std::isfinite(dz_box)) return fail(c, YTF_EINVAL, "Δz_box = %g must be ≥ 0", dz_box);
  for (int p = 0; p < n_pairs; ++p) {
    const double zp = source_layer == 1 ? ds - zsrc[p] : zsrc[p];
    const double tol = 1.4901161193847656e-08;
    if (!(zp >= -tol && zp <= ds + tol))
      return fail(c, YTF_EINVAL, "amplitude_at: zsrc_local=%g outside [0,%g]", zsrc[p], ds);
    // _check_box_in_layer (mod_ClosedFormField.jl:93-98)
    const double bt = tol * std::max(ds, 1.0);
    if (dz_box > 0.0 && !(zp - dz_box / 2 >= -bt && zp + dz_box / 2 <= ds + bt))
      return fail(c, YTF_EINVAL, "box [%g, %g] (cavity frame) does not fit inside its layer [0, %g]", zp - dz_box / 2,
                  zp + dz_box / 2, ds);
  }
  if (n_nodes == 0) return YTF_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t lb = sizeof(ytf_layer) * n_layers, zb = sizeof(double) * (size_t)n_pairs;
  std::vector<char> hb(lb + 2 * zb);
  std::memcpy(hb.data(), layers, lb);
  if (n_pairs) {
    std::memcpy(hb.data() + lb, zsrc, zb);
    std::memcpy(hb.data() + lb + zb, zobs, zb);
  }
  if (int rc = set_small(c, hb)) return rc;
  int rc_redo = ensure_redo(c, n_nodes);
  if (rc_redo) return rc_redo;
  char *sm = (char *)c->small;
  ytf::NodeParams &P = *P_out;
  P.L = (const LayerDev *)sm;
  P.nl = n_layers;
  P.omega = omega;
  P.kx = kx_dev;
  P.ky = ky_dev;
  P.nn = n_nodes;
  P.src = source_layer - 1;
  P.obs = obs_layer - 1;
  P.npairs = n_pairs;
  P.terms = terms;
  P.dz = dz_box;
  P.ez_local = ez_local ? 1 : 0;
  P.zsrc = (const double *)(sm + lb);
  P.zobs = (const double *)(sm + lb + zb);
  P.ghat = (cd *)ghat_dev;
  P.Qsrc = cache ? (cd *)cache->Qsrc_dev : nullptr;
  P.sf = cache ? (cd *)cache->sf_dev : nullptr;
  P.sb = cache ? (cd *)cache->sb_dev : nullptr;
  P.T = cache ? (cd *)cache->T_dev : nullptr;
  P.basis = cache ? (cd *)cache->basis_dev : nullptr;
  P.flags = cache ? (unsigned char *)cache->flags_dev : nullptr;
  if ((P.sf == nullptr) != (P.sb == nullptr)) return fail(c, YTF_EINVAL, "ytf_green_nodes: sf and sb go together");
  P.needs_general = stack_needs_general(layers, n_layers);
  P.redo = (int *)c->redo;
  return YTF_OK;
}

// One node solve of P.nn nodes on stream s (one wave of 64 nodes per block).
static hipError_t launch_node_solve(ytf_ctx *c, const ytf::NodeParams &P) {
  int64_t blocks = (P.nn + 63) / 64;
  if (blocks > 1 << 20) blocks = 1 << 20;
  // The node kernels run one wave per SIMD.  When the nodes fill less than
  // one round of the GPU's SIMDs and there are several depth pairs, the pairs
  // are split over the idle SIMDs (each group repeats the node solve, and the
  // per-pair g~ work shrinks by the group count); the groups fill one round.
  int groups = 1;
  if (P.ghat && P.npairs > 1 && !P.rows) {
    const int64_t slots = 4 * (int64_t)std::max(c->n_cu, 1);
    groups = (int)std::max<int64_t>(1, std::min<int64_t>(slots / blocks, P.npairs));
    const int ppg = (P.npairs + groups - 1) / groups;
    groups = (P.npairs + ppg - 1) / ppg;  // no empty group
  }
  return (P.src == 0 && P.obs == 0) ? ytf::launch_node_top(P, (int)blocks, c->stream, groups)
                                    : ytf::launch_node_general(P, (int)blocks, c->stream, groups);
}

int green_nodes_locked(ytf_ctx *c, const ytf_layer *layers, int n_layers, double omega, const double *kx_dev,
                       const double *ky_dev, int64_t n_nodes, int source_layer, int obs_layer, int n_pairs,
                       const double *zsrc, const double *zobs, double *ghat_dev, const ytf_node_cache *cache,
                       int terms, double dz_box, int ez_local) {
  ytf::NodeParams P;
  int rc = node_params_locked(c, layers, n_layers, omega, kx_dev, ky_dev, n_nodes, source_layer, obs_layer, n_pairs,
                              zsrc, zobs, ghat_dev, cache, terms, dz_box, ez_local, &P);
  if (rc || P.nn == 0) return rc;
  hipError_t e = timed(c, "node_solve", [&] { return launch_node_solve(c, P); });
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_green_nodes: launch failed: %s", hipGetErrorString(e));
  return YTF_OK;
}

extern "C" {

int ytf_pole_candidates(ytf_ctx *c, int n_dirs, int64_t n, const double *R_dev, const double *ks_dev, int W,
                        double log_prom, double log_floor_rel, double log_floor_abs, int max_out, int64_t *key_dev,
                        double *val_dev, int64_t *n_out) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_pole_candidates: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n_dirs < 0 || n < 0 || W < 0 || max_out < 0 || !n_out || (n_dirs > 0 && n > 0 && (!R_dev || !ks_dev)) ||
      (max_out > 0 && (!key_dev || !val_dev)) || (int64_t)n_dirs * n > INT64_MAX / 2)
    return fail(c, YTF_EINVAL, "ytf_pole_candidates: bad arguments");
  *n_out = 0;
  if (n_dirs == 0 || n < 3) return YTF_OK;
  HIPCHK(c, hipSetDevice(c->device));
  int rc = ensure_pinned(c);
  if (rc) return rc;
  const size_t yb = align256(sizeof(double) * (size_t)n_dirs * (size_t)n);
  rc = ensure_buf(c, &c->aux, &c->aux_bytes, align256(sizeof(int)) + align256(sizeof(double) * (size_t)n_dirs) + yb);
  if (rc) return rc;
  int *count = (int *)c->aux;
  double *ymax = (double *)((char *)c->aux + align256(sizeof(int)));
  double *Y = (double *)((char *)ymax + align256(sizeof(double) * (size_t)n_dirs));
  hipError_t e = ytf::launch_mlp_zero(1, (float *)count, c->stream);  // zero bits = integer 0
  if (e == hipSuccess)
    e = timed(c, "pole_candidates", [&] {
      return ytf::launch_pole_candidates(n_dirs, n, W, log_prom, log_floor_rel, log_floor_abs, R_dev, ks_dev, Y, ymax,
                                         max_out, count, key_dev, val_dev, c->stream);
    });
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_pole_candidates: launch failed: %s", hipGetErrorString(e));
  HIPCHK(c, hipMemcpyAsync(c->pinned, count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *n_out = c->pinned[0];
  return YTF_OK;
}

}  // extern "C"

// Validate the stack of a cavity-response call and send it to the device
// (when it differs from what the device copy holds).
static int response_stack_locked(ytf_ctx *c, const char *who, const ytf_layer *layers, int n_layers, double omega,
                                 int source_layer) {
  if (!layers || n_layers < 2) return fail(c, YTF_EINVAL, "%s: need at least 2 layers", who);
  if (source_layer < 1 || source_layer > n_layers - 1)
    return fail(c, YTF_EINVAL, "%s: source_layer=%d must be in 1:%d", who, source_layer, n_layers - 1);
  if (!(omega > 0) || !std::isfinite(omega)) return fail(c, YTF_EINVAL, "%s: omega must be > 0", who);
  for (int l = 0; l < n_layers; ++l)
    if (layers[l].eps_kind < 0 || layers[l].eps_kind > 3 || layers[l].mu_kind < 0 || layers[l].mu_kind > 1)
      return fail(c, YTF_EINVAL, "%s: layer %d has an unknown material kind", who, l + 1);
  HIPCHK(c, hipSetDevice(c->device));
  const size_t lb = sizeof(ytf_layer) * n_layers;
  std::vector<char> hb(lb);
  std::memcpy(hb.data(), layers, lb);
  return set_small(c, hb);
}

extern "C" {

int ytf_cavity_response(ytf_ctx *c, const ytf_layer *layers, int n_layers, double omega, int source_layer,
                        const double *kx_dev, const double *ky_dev, int64_t n_nodes, double *out_dev) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_cavity_response: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n_nodes < 0 || (n_nodes > 0 && (!kx_dev || !ky_dev || !out_dev)))
    return fail(c, YTF_EINVAL, "ytf_cavity_response: bad nodes");
  int rc = response_stack_locked(c, "ytf_cavity_response", layers, n_layers, omega, source_layer);
  if (rc || n_nodes == 0) return rc;
  if (n_nodes > INT32_MAX) return fail(c, YTF_EINVAL, "ytf_cavity_response: at most 2^31 - 1 nodes per call");
  rc = ensure_redo(c, n_nodes);
  if (rc) return rc;
  const bool general = stack_needs_general(layers, n_layers);
  hipError_t e = timed(c, "cavity_response", [&] {
    return ytf::launch_cavity_response((const LayerDev *)c->small, n_layers, source_layer, omega, kx_dev, ky_dev,
                                       n_nodes, out_dev, general, (int *)c->redo, c->stream);
  });
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_cavity_response: launch failed: %s", hipGetErrorString(e));
  return YTF_OK;
}

int ytf_pole_refine(ytf_ctx *c, const ytf_layer *layers, int n_layers, double omega, int source_layer,
                    int64_t n_cand, const double *cand_dev, double *out_dev, int64_t *n_evals) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_pole_refine: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n_cand < 0 || n_cand > INT32_MAX || (n_cand > 0 && (!cand_dev || !out_dev)))
    return fail(c, YTF_EINVAL, "ytf_pole_refine: bad candidates");
  if (n_evals) *n_evals = 0;
  int rc = response_stack_locked(c, "ytf_pole_refine", layers, n_layers, omega, source_layer);
  if (rc || n_cand == 0) return rc;
  const bool general = stack_needs_general(layers, n_layers);
  unsigned long long *evals = nullptr;
  if (n_evals) {
    if ((rc = ensure_pinned(c)) || (rc = ensure_buf(c, &c->aux, &c->aux_bytes, 256))) return rc;
    evals = (unsigned long long *)c->aux;
    hipError_t z = ytf::launch_mlp_zero(2, (float *)evals, c->stream);  // zero bits = integer 0
    if (z != hipSuccess) return fail(c, YTF_EHIP, "ytf_pole_refine: launch failed: %s", hipGetErrorString(z));
  }
  hipError_t e = timed(c, "pole_refine", [&] {
    return ytf::launch_pole_refine((const LayerDev *)c->small, n_layers, source_layer, omega, n_cand, cand_dev,
                                   out_dev, evals, general, c->n_cu, c->stream);
  });
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_pole_refine: launch failed: %s", hipGetErrorString(e));
  if (n_evals) {
    HIPCHK(c, hipMemcpyAsync(c->pinned, evals, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *n_evals = (int64_t) * (unsigned long long *)c->pinned;
  }
  return YTF_OK;
}

int ytf_spectral_apply(ytf_ctx *c, int nx, int ny, const double *ghat_dev, const double *J_dev, double *E_dev,
                       double *work_dev, double scale) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_spectral_apply: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!ytf::fft_length_ok(nx) || !ytf::fft_length_ok(ny))
    return fail(c, YTF_EINVAL,
                "ytf_spectral_apply: nx=%d, ny=%d unsupported (powers of two 8..2048, any other size 1..1024)", nx, ny);
  if (!ghat_dev || !J_dev || !E_dev || !work_dev) return fail(c, YTF_EINVAL, "ytf_spectral_apply: NULL buffer");
  if ((const void *)E_dev == (const void *)J_dev) return fail(c, YTF_EINVAL, "ytf_spectral_apply: E aliases J");
  HIPCHK(c, hipSetDevice(c->device));
  cd *twx = nullptr, *twy = nullptr;
  int rc = twiddles(c, nx, &twx);
  if (rc) return rc;
  rc = twiddles(c, ny, &twy);
  if (rc) return rc;
  hipError_t e = timed(c, "spectral_apply", [&] {
    return ytf::launch_spectral_apply(nx, ny, (const cd *)ghat_dev, (const cd *)J_dev, (cd *)E_dev, (cd *)work_dev,
                                      twx, twy, scale, c->stream);
  });
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_spectral_apply: launch failed: %s", hipGetErrorString(e));
  return YTF_OK;
}

int ytf_ctx_set_timing(ytf_ctx *c, int on) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_ctx_set_timing: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  int rc = drain_marks(c);
  if (rc) return rc;
  c->totals.clear();
  c->timing = on != 0;
  c->timing_only.clear();
  return YTF_OK;
}

int ytf_ctx_set_timing_only(ytf_ctx *c, const char *name) {
  if (!c || !name || !*name) return fail(c, YTF_EINVAL, "ytf_ctx_set_timing_only: NULL or empty argument");
  std::lock_guard<std::mutex> lk(c->mu);
  int rc = drain_marks(c);
  if (rc) return rc;
  c->totals.clear();
  c->timing = true;
  c->timing_only.clear();
  for (const char *p = name;;) {
    const char *q = std::strchr(p, ',');
    const std::string t = q ? std::string(p, q - p) : std::string(p);
    if (!t.empty()) c->timing_only.push_back(t);
    if (!q) break;
    p = q + 1;
  }
  if (c->timing_only.empty()) c->timing = false;
  return YTF_OK;
}

int ytf_ctx_kernel_time(ytf_ctx *c, const char *name, double *total_ms, int64_t *launches) {
  if (!c || !name || !total_ms || !launches) return fail(c, YTF_EINVAL, "ytf_ctx_kernel_time: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  int rc = drain_marks(c);
  if (rc) return rc;
  auto it = c->totals.find(name);
  *total_ms = it == c->totals.end() ? 0.0 : it->second.first;
  *launches = it == c->totals.end() ? 0 : it->second.second;
  return YTF_OK;
}

// ---------------------------------------------------------------------------
// Uneven batches of stacks (batch.hip)
// ---------------------------------------------------------------------------
struct ytf_batch {
  int device = 0;
  int n_stacks = 0;
  int64_t ntot = 0;
  double omega = 0.0;
  bool general = false;
  std::vector<int> l0, nl;
  std::vector<int64_t> n0, nn;
  // device tables
  void *dev = nullptr;  // layers | l0 | nl | n0 | nn | items_stack | items_first | counter
  LayerDev *L = nullptr;
  int *d_l0 = nullptr, *d_nl = nullptr, *d_is = nullptr, *d_if = nullptr;
  int64_t *d_n0 = nullptr, *d_nn = nullptr;
  unsigned int *counter = nullptr;
  int *redo = nullptr;  // redo marks of the nodes whose interface solve was exactly singular
  int n_items = 0;
  // launch configuration
  int nodes_per_item = 128, waves_per_block = 1, blocks = 0;
  int items_for = 0;  // nodes_per_item the item table was built for
  // captured graph of (counter reset + persistent kernel + redo) for one argument set
  hipGraphExec_t exec = nullptr;
  std::vector<const void *> exec_key;
  std::string signature;
};


// Work items: runs of nodes_per_item nodes of one stack, heaviest stacks first.
static int batch_items(ytf_ctx *c, ytf_batch *b) {
  if (b->items_for == b->nodes_per_item) return YTF_OK;
  std::vector<int> order(b->n_stacks);
  for (int s = 0; s < b->n_stacks; ++s) order[s] = s;
  std::stable_sort(order.begin(), order.end(), [&](int a, int q) { return b->nl[a] > b->nl[q]; });
  std::vector<int> is, ifirst;
  for (int s : order)
    for (int64_t f = 0; f < b->nn[s]; f += b->nodes_per_item) {
      is.push_back(s);
      ifirst.push_back((int)f);
    }
  b->n_items = (int)is.size();
  // the item tables were sized for 64-node items
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (b->n_items) {
    HIPCHK(c, hipMemcpy(b->d_is, is.data(), sizeof(int) * is.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_if, ifirst.data(), sizeof(int) * ifirst.size(), hipMemcpyHostToDevice));
  }
  b->items_for = b->nodes_per_item;
  return YTF_OK;
}

static ytf::BatchLaunch batch_launch(ytf_batch *b, const double *kx, const double *ky, double *S, double *V1,
                                     double *Q1, double *VN, double *QN, uint8_t *flags) {
  ytf::BatchLaunch B;
  B.L = b->L;
  B.stack_l0 = b->d_l0;
  B.stack_nl = b->d_nl;
  B.stack_n0 = b->d_n0;
  B.stack_nn = b->d_nn;
  B.item_stack = b->d_is;
  B.item_first = b->d_if;
  B.n_items = b->n_items;
  B.nodes_per_item = b->nodes_per_item;
  B.omega = b->omega;
  B.kx = kx;
  B.ky = ky;
  B.S = (cd *)S;
  B.V1 = (cd *)V1;
  B.Q1 = (cd *)Q1;
  B.VN = (cd *)VN;
  B.QN = (cd *)QN;
  B.flags = flags;
  B.ntot = b->ntot;
  B.counter = b->counter;
  B.redo = b->redo;
  B.n_stacks = b->n_stacks;
  B.general = b->general;
  B.waves_per_block = b->waves_per_block;
  B.blocks = b->blocks;
  return B;
}

static int device_cus(ytf_ctx *c, int *cus) {
  HIPCHK(c, hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, c->device));
  return YTF_OK;
}

static int batch_check_cfg(ytf_ctx *c, int npi, int wpb, int blocks) {
  if (npi < 64 || npi > 4096 || (npi % 64) != 0)
    return fail(c, YTF_EINVAL, "ytf_batch: nodes_per_item=%d must be a multiple of 64 in 64..4096", npi);
  if (wpb != 1 && wpb != 2 && wpb != 4) return fail(c, YTF_EINVAL, "ytf_batch: waves_per_block=%d not in {1,2,4}", wpb);
  if (blocks < 1 || blocks > 65536) return fail(c, YTF_EINVAL, "ytf_batch: blocks=%d out of range 1..65536", blocks);
  return YTF_OK;
}

int ytf_batch_create(ytf_ctx *c, const ytf_layer *layers, const int32_t *layer_counts, const int64_t *node_counts,
                     int n_stacks, double omega, ytf_batch **out) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_batch_create: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!out || !layers || !layer_counts || !node_counts || n_stacks < 1)
    return fail(c, YTF_EINVAL, "ytf_batch_create: bad arguments");
  if (!(omega > 0) || !std::isfinite(omega)) return fail(c, YTF_EINVAL, "ytf_batch_create: omega must be > 0");
  ytf_batch *b = new ytf_batch();
  b->device = c->device;
  b->n_stacks = n_stacks;
  b->omega = omega;
  int nlt = 0;
  int64_t work = 0;
  for (int s = 0; s < n_stacks; ++s) {
    if (layer_counts[s] < 2 || node_counts[s] < 0 || node_counts[s] > ((int64_t)1 << 31) - 4096) {
      delete b;
      return fail(c, YTF_EINVAL, "ytf_batch_create: stack %d has %d layers and %lld nodes (need >= 2 layers)", s,
                  layer_counts[s], (long long)node_counts[s]);
    }
    b->l0.push_back(nlt);
    b->nl.push_back(layer_counts[s]);
    b->n0.push_back(b->ntot);
    b->nn.push_back(node_counts[s]);
    nlt += layer_counts[s];
    b->ntot += node_counts[s];
    work += node_counts[s] * layer_counts[s];
  }
  for (int l = 0; l < nlt; ++l) {
    const ytf_layer &Ly = layers[l];
    if (Ly.eps_kind < 0 || Ly.eps_kind > 3 || Ly.mu_kind < 0 || Ly.mu_kind > 1) {
      delete b;
      return fail(c, YTF_EINVAL, "ytf_batch_create: layer %d has an unknown material kind", l);
    }
    if ((Ly.eps_kind == 0 || Ly.eps_kind == 1) && Ly.mu_kind == 0) continue;
    const int idx[4] = {2, 5, 6, 7};
    for (int t = 0; t < 4; ++t)
      if (Ly.eps[2 * idx[t]] != 0.0 || Ly.eps[2 * idx[t] + 1] != 0.0 || Ly.mu[2 * idx[t]] != 0.0 ||
          Ly.mu[2 * idx[t] + 1] != 0.0)
        b->general = true;
  }
  int64_t max_items = 0;
  for (int s = 0; s < n_stacks; ++s) max_items += (b->nn[s] + 63) / 64;
  const size_t lb = align256(sizeof(ytf_layer) * nlt), ib = align256(sizeof(int) * n_stacks),
               jb = align256(sizeof(int64_t) * n_stacks), tb = align256(sizeof(int) * (size_t)(max_items + 1));
  // redo marks: two words and a byte per node
  const size_t rb = align256(2 * sizeof(int) + (size_t)b->ntot);
  if (hipSetDevice(c->device) != hipSuccess ||
      hipMalloc(&b->dev, lb + 2 * ib + 2 * jb + 2 * tb + 256 + rb) != hipSuccess) {
    delete b;
    return fail(c, YTF_EHIP, "ytf_batch_create: device allocation failed");
  }
  char *d = (char *)b->dev;
  b->L = (LayerDev *)d;
  b->d_l0 = (int *)(d + lb);
  b->d_nl = (int *)(d + lb + ib);
  b->d_n0 = (int64_t *)(d + lb + 2 * ib);
  b->d_nn = (int64_t *)(d + lb + 2 * ib + jb);
  b->d_is = (int *)(d + lb + 2 * ib + 2 * jb);
  b->d_if = (int *)(d + lb + 2 * ib + 2 * jb + tb);
  b->counter = (unsigned int *)(d + lb + 2 * ib + 2 * jb + 2 * tb);
  b->redo = (int *)(d + lb + 2 * ib + 2 * jb + 2 * tb + 256);
  hipError_t e = hipMemcpy(b->L, layers, sizeof(ytf_layer) * nlt, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(b->redo, 0, rb);
  if (e == hipSuccess) e = hipMemcpy(b->d_l0, b->l0.data(), sizeof(int) * n_stacks, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_nl, b->nl.data(), sizeof(int) * n_stacks, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_n0, b->n0.data(), sizeof(int64_t) * n_stacks, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_nn, b->nn.data(), sizeof(int64_t) * n_stacks, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(b->dev);
    delete b;
    return fail(c, YTF_EHIP, "ytf_batch_create: upload failed: %s", hipGetErrorString(e));
  }
  int cus = 0;
  if (device_cus(c, &cus)) {
    hipFree(b->dev);
    delete b;
    return YTF_EHIP;
  }
  b->blocks = 4 * cus;  // one wave per SIMD: the kernel's register budget allows no more
  std::ostringstream sig;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) sig << prop.gcnArchName << "/" << cus << "cu";
  sig << "/s" << n_stacks << "/n" << b->ntot << "/w" << work << "/g" << (int)b->general;
  b->signature = sig.str();
  for (auto &ch : b->signature)
    if (ch == ' ') ch = '_';
  *out = b;
  return YTF_OK;
}

int ytf_batch_destroy(ytf_ctx *c, ytf_batch *b) {
  if (!b) return YTF_OK;
  if (c) {
    std::lock_guard<std::mutex> lk(c->mu);
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
  }
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->dev) hipFree(b->dev);
  delete b;
  return YTF_OK;
}

int ytf_batch_set_config(ytf_ctx *c, ytf_batch *b, int nodes_per_item, int waves_per_block, int blocks) {
  if (!c || !b) return fail(c, YTF_EINVAL, "ytf_batch_set_config: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  int rc = batch_check_cfg(c, nodes_per_item, waves_per_block, blocks);
  if (rc) return rc;
  b->nodes_per_item = nodes_per_item;
  b->waves_per_block = waves_per_block;
  b->blocks = blocks;
  return YTF_OK;
}

int ytf_batch_get_config(ytf_ctx *c, const ytf_batch *b, int *nodes_per_item, int *waves_per_block, int *blocks) {
  if (!c || !b || !nodes_per_item || !waves_per_block || !blocks)
    return fail(c, YTF_EINVAL, "ytf_batch_get_config: NULL argument");
  *nodes_per_item = b->nodes_per_item;
  *waves_per_block = b->waves_per_block;
  *blocks = b->blocks;
  return YTF_OK;
}

static int batch_solve_locked(ytf_ctx *c, ytf_batch *b, const double *kx, const double *ky, double *S, double *V1,
                              double *Q1, double *VN, double *QN, uint8_t *flags, bool use_graph) {
  if (b->ntot == 0) return YTF_OK;
  if (!kx || !ky || !S) return fail(c, YTF_EINVAL, "ytf_batch_solve: kx, ky and S are required");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = batch_items(c, b);
  if (rc) return rc;
  const ytf::BatchLaunch B = batch_launch(b, kx, ky, S, V1, Q1, VN, QN, flags);
  if (!use_graph) {
    hipError_t e = timed(c, "batch_solve", [&] { return ytf::launch_batch(B, c->stream); });
    if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_batch_solve: launch failed: %s", hipGetErrorString(e));
    return YTF_OK;
  }
  // one graph per argument set: (counter reset, persistent kernel, redo)
  std::vector<const void *> key = {kx, ky, S, V1, Q1, VN, QN, flags, (const void *)(intptr_t)b->nodes_per_item,
                                   (const void *)(intptr_t)b->waves_per_block, (const void *)(intptr_t)b->blocks};
  if (!b->exec || key != b->exec_key) {
    if (b->exec) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipGraphExecDestroy(b->exec));
      b->exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->own, hipStreamCaptureModeThreadLocal));
    hipError_t e = ytf::launch_batch(B, c->own);
    hipError_t e2 = hipStreamEndCapture(c->own, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph) hipGraphDestroy(graph);
      return fail(c, YTF_EHIP, "ytf_batch_solve: graph capture failed: %s",
                  hipGetErrorString(e != hipSuccess ? e : e2));
    }
    hipError_t e3 = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e3 != hipSuccess) {
      b->exec = nullptr;
      return fail(c, YTF_EHIP, "ytf_batch_solve: graph instantiation failed: %s", hipGetErrorString(e3));
    }
    b->exec_key = key;
  }
  hipError_t e = timed(c, "batch_solve", [&] { return hipGraphLaunch(b->exec, c->stream); });
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_batch_solve: graph launch failed: %s", hipGetErrorString(e));
  return YTF_OK;
}

int ytf_batch_solve(ytf_ctx *c, ytf_batch *b, const double *kx_dev, const double *ky_dev, double *S_dev,
                    double *V1_dev, double *Q1_dev, double *VN_dev, double *QN_dev, uint8_t *flags_dev) {
  if (!c || !b) return fail(c, YTF_EINVAL, "ytf_batch_solve: NULL ctx or batch");
  std::lock_guard<std::mutex> lk(c->mu);
  if (b->device != c->device) return fail(c, YTF_EINVAL, "ytf_batch_solve: batch belongs to device %d", b->device);
  return batch_solve_locked(c, b, kx_dev, ky_dev, S_dev, V1_dev, Q1_dev, VN_dev, QN_dev, flags_dev, true);
}

// Autotuner: time every candidate (nodes per item x waves per block x grid
// size) on this batch's own inputs, keep the fastest.  With a tune-DB file, a
// stored choice for the same batch signature is reused without timing, and a
// new choice is appended.
int ytf_batch_tune(ytf_ctx *c, ytf_batch *b, const double *kx_dev, const double *ky_dev, double *S_dev,
                   uint8_t *flags_dev, const char *tune_db, char *report, int report_len) {
  if (!c || !b) return fail(c, YTF_EINVAL, "ytf_batch_tune: NULL ctx or batch");
  std::lock_guard<std::mutex> lk(c->mu);
  auto say = [&](const std::string &m) {
    if (report && report_len > 0) {
      std::strncpy(report, m.c_str(), report_len - 1);
      report[report_len - 1] = 0;
    }
  };
  if (tune_db && *tune_db) {
    std::ifstream f(tune_db);
    std::string line;
    while (std::getline(f, line)) {
      std::istringstream ls(line);
      std::string sig;
      int npi, wpb, blk;
      double ms;
      if (ls >> sig >> npi >> wpb >> blk >> ms && sig == b->signature && !batch_check_cfg(c, npi, wpb, blk)) {
        b->nodes_per_item = npi;
        b->waves_per_block = wpb;
        b->blocks = blk;
        say("db " + b->signature + " " + std::to_string(npi) + " " + std::to_string(wpb) + " " + std::to_string(blk));
        return YTF_OK;
      }
    }
  }
  if (b->ntot == 0) return YTF_OK;
  int cus = 0;
  int rc = device_cus(c, &cus);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  double best = 1e300;
  int bn = b->nodes_per_item, bw = b->waves_per_block, bb = b->blocks;
  const int npis[4] = {64, 128, 256, 512};
  const int wpbs[3] = {1, 2, 4};
  std::ostringstream log;
  for (int npi : npis)
    for (int wpb : wpbs)
      for (int over = 1; over <= 2; ++over) {
        b->nodes_per_item = npi;
        b->waves_per_block = wpb;
        b->blocks = over * 4 * cus / wpb;
        double ts[3];
        for (int r = 0; r < 4; ++r) {
          hipEventRecord(e0, c->stream);
          rc = batch_solve_locked(c, b, kx_dev, ky_dev, S_dev, nullptr, nullptr, nullptr, nullptr, flags_dev, false);
          if (rc) {
            hipEventDestroy(e0);
            hipEventDestroy(e1);
            return rc;
          }
          hipEventRecord(e1, c->stream);
          hipEventSynchronize(e1);
          float ms = 0;
          hipEventElapsedTime(&ms, e0, e1);
          if (r) ts[r - 1] = ms;
        }
        std::sort(ts, ts + 3);
        log << npi << "/" << wpb << "/" << b->blocks << ":" << ts[1] << " ";
        if (ts[1] < best) {
          best = ts[1];
          bn = npi;
          bw = wpb;
          bb = b->blocks;
        }
      }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  b->nodes_per_item = bn;
  b->waves_per_block = bw;
  b->blocks = bb;
  if (tune_db && *tune_db) {
    std::ofstream f(tune_db, std::ios::app);
    if (f) f << b->signature << " " << bn << " " << bw << " " << bb << " " << best << "\n";
  }
  std::ostringstream m;
  m << "tuned " << b->signature << " best " << bn << " " << bw << " " << bb << " " << best << " ms; " << log.str();
  say(m.str());
  return YTF_OK;
}

// Radial synthesis into R (device, [np][9][nr] complex); the caller holds the lock.

// ---------------------------------------------------------------------------
// Surrogate network (mlp.hip)
// ---------------------------------------------------------------------------
struct ytf_mlp {
  int device = 0;
  int nin = 0, width = 0, depth = 0, nout = 0, nl = 0;
  int kzf = 0;  // two light-line features follow the inputs
  std::vector<int> K, N, Kp, Np;       // real and padded (to 64) layer shapes
  std::vector<size_t> offW, offb;      // offsets in the padded parameter array
  std::vector<size_t> foffW, foffb;    // offsets in the flat (unpadded) layout
  size_t np = 0, nflat = 0;
  float *P = nullptr, *Mo = nullptr, *Vo = nullptr, *G = nullptr;  // FP32 master, Adam moments, gradient
  __bf16 *P16 = nullptr;                                            // bf16 copy of P
  __bf16 *P16pk = nullptr;  // weights of P16 in matrix-core fragment order (the fused forward)
  uint64_t version = 1, packed_version = 0;  // P16 changes / the version P16pk was packed from
  float *norm = nullptr;  // in_shift[nin] in_scale[nin] out_shift[nout] out_scale[nout]
  float *loss = nullptr;
  long step = 0;
  // activation workspace
  void *act = nullptr;
  size_t act_bytes = 0;
  // k_rho windows of the surrogate's pole flag (lo, hi pairs), device
  double *pole_win = nullptr;
  int n_pole_win = 0;
};

static size_t ru64(size_t x) { return (x + 63) / 64 * 64; }

static int mlp_free(ytf_mlp *m) {
  void *ps[] = {m->P, m->Mo, m->Vo, m->G, m->P16, m->P16pk, m->norm, m->loss, m->act, m->pole_win};
  for (void *q : ps)
    if (q) hipFree(q);
  delete m;
  return YTF_OK;
}

int ytf_mlp_create(ytf_ctx *c, int n_in, int width, int depth, int n_out, int flags, uint64_t seed, ytf_mlp **out) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_mlp_create: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  const int kzf = (flags & YTF_MLP_KZ_FEATURES) ? 1 : 0;
  if (!out || n_in < 1 || n_in + 2 * kzf > 64 || (kzf && n_in < 2) || width < 64 || width > 4096 || width % 64 ||
      depth < 1 || depth > 32 || n_out < 1 || n_out > 4096 || (flags & ~YTF_MLP_KZ_FEATURES))
    return fail(c, YTF_EINVAL,
                "ytf_mlp_create: bad shape (n_in=%d, width=%d (multiple of 64), depth=%d, n_out=%d, flags=%d)", n_in,
                width, depth, n_out, flags);
  ytf_mlp *m = new ytf_mlp();
  m->device = c->device;
  m->kzf = kzf;
  m->nin = n_in;
  m->width = width;
  m->depth = depth;
  m->nout = n_out;
  m->nl = depth + 1;
  for (int l = 0; l < m->nl; ++l) {
    const int k = l == 0 ? n_in + 2 * kzf : width, n = l == m->nl - 1 ? n_out : width;
    m->K.push_back(k);
    m->N.push_back(n);
    m->Kp.push_back((int)ru64(k));
    m->Np.push_back((int)ru64(n));
    m->offW.push_back(m->np);
    m->np += (size_t)m->Np[l] * m->Kp[l];
    m->offb.push_back(m->np);
    m->np += m->Np[l];
    m->foffW.push_back(m->nflat);
    m->nflat += (size_t)n * k;
    m->foffb.push_back(m->nflat);
    m->nflat += n;
  }
  hipError_t e = hipSetDevice(c->device);
  const size_t fb = sizeof(float) * m->np;
  if (e == hipSuccess) e = hipMalloc(&m->P, fb);
  if (e == hipSuccess) e = hipMalloc(&m->Mo, fb);
  if (e == hipSuccess) e = hipMalloc(&m->Vo, fb);
  if (e == hipSuccess) e = hipMalloc(&m->G, fb);
  if (e == hipSuccess) e = hipMalloc(&m->P16, sizeof(__bf16) * m->np);
  if (e == hipSuccess) e = hipMalloc(&m->P16pk, sizeof(__bf16) * m->np);
  if (e == hipSuccess) e = hipMalloc(&m->norm, sizeof(float) * 2 * (n_in + n_out));
  if (e == hipSuccess) e = hipMalloc(&m->loss, sizeof(float));
  if (e != hipSuccess) {
    mlp_free(m);
    return fail(c, YTF_EHIP, "ytf_mlp_create: allocation failed: %s", hipGetErrorString(e));
  }
  // Xavier-uniform weights, zero biases, zero padding; identity normalisation
  std::vector<float> h(m->np, 0.0f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < m->nl; ++l) {
    const float a = std::sqrt(6.0f / (float)(m->K[l] + m->N[l]));
    std::uniform_real_distribution<float> U(-a, a);
    for (int i = 0; i < m->N[l]; ++i)
      for (int j = 0; j < m->K[l]; ++j) h[m->offW[l] + (size_t)i * m->Kp[l] + j] = U(rng);
  }
  std::vector<float> nrm(2 * (n_in + n_out), 0.0f);
  for (int i = 0; i < n_in; ++i) nrm[n_in + i] = 1.0f;
  for (int i = 0; i < n_out; ++i) nrm[2 * n_in + n_out + i] = 1.0f;
  std::vector<float> zero(m->np, 0.0f);
  e = hipMemcpy(m->P, h.data(), fb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->Mo, zero.data(), fb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->Vo, zero.data(), fb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->G, zero.data(), fb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->norm, nrm.data(), sizeof(float) * nrm.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = ytf::launch_mlp_to_bf16((int64_t)m->np, m->P, m->P16, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    mlp_free(m);
    return fail(c, YTF_EHIP, "ytf_mlp_create: initialisation failed: %s", hipGetErrorString(e));
  }
  *out = m;
  return YTF_OK;
}

int ytf_mlp_destroy(ytf_ctx *c, ytf_mlp *m) {
  if (!m) return YTF_OK;
  if (c) {
    std::lock_guard<std::mutex> lk(c->mu);
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
  }
  return mlp_free(m);
}

int64_t ytf_mlp_num_params(const ytf_mlp *m) { return m ? (int64_t)m->nflat : 0; }

// flat (unpadded) <-> padded parameter layouts, on the host
int ytf_mlp_get_params(ytf_ctx *c, ytf_mlp *m, float *host) {
  if (!c || !m || !host) return fail(c, YTF_EINVAL, "ytf_mlp_get_params: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<float> h(m->np);
  HIPCHK(c, hipMemcpy(h.data(), m->P, sizeof(float) * m->np, hipMemcpyDeviceToHost));
  for (int l = 0; l < m->nl; ++l) {
    for (int i = 0; i < m->N[l]; ++i)
      for (int j = 0; j < m->K[l]; ++j) host[m->foffW[l] + (size_t)i * m->K[l] + j] = h[m->offW[l] + (size_t)i * m->Kp[l] + j];
    for (int i = 0; i < m->N[l]; ++i) host[m->foffb[l] + i] = h[m->offb[l] + i];
  }
  return YTF_OK;
}

int ytf_mlp_set_params(ytf_ctx *c, ytf_mlp *m, const float *host) {
  if (!c || !m || !host) return fail(c, YTF_EINVAL, "ytf_mlp_set_params: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<float> h(m->np, 0.0f), zero(m->np, 0.0f);
  for (int l = 0; l < m->nl; ++l) {
    for (int i = 0; i < m->N[l]; ++i)
      for (int j = 0; j < m->K[l]; ++j) h[m->offW[l] + (size_t)i * m->Kp[l] + j] = host[m->foffW[l] + (size_t)i * m->K[l] + j];
    for (int i = 0; i < m->N[l]; ++i) h[m->offb[l] + i] = host[m->foffb[l] + i];
  }
  HIPCHK(c, hipMemcpy(m->P, h.data(), sizeof(float) * m->np, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(m->Mo, zero.data(), sizeof(float) * m->np, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(m->Vo, zero.data(), sizeof(float) * m->np, hipMemcpyHostToDevice));
  m->step = 0;
  m->version += 1;
  hipError_t e = ytf::launch_mlp_to_bf16((int64_t)m->np, m->P, m->P16, c->stream);
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_mlp_set_params: %s", hipGetErrorString(e));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return YTF_OK;
}

int ytf_mlp_set_normalization(ytf_ctx *c, ytf_mlp *m, const float *in_shift, const float *in_scale,
                              const float *out_shift, const float *out_scale) {
  if (!c || !m || !in_shift || !in_scale || !out_shift || !out_scale)
    return fail(c, YTF_EINVAL, "ytf_mlp_set_normalization: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  std::vector<float> h;
  h.insert(h.end(), in_shift, in_shift + m->nin);
  h.insert(h.end(), in_scale, in_scale + m->nin);
  h.insert(h.end(), out_shift, out_shift + m->nout);
  h.insert(h.end(), out_scale, out_scale + m->nout);
  for (int i = 0; i < m->nout; ++i)
    if (!(out_scale[i] != 0.0f) || !std::isfinite(out_scale[i]))
      return fail(c, YTF_EINVAL, "ytf_mlp_set_normalization: out_scale[%d] must be finite and nonzero", i);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(m->norm, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
  return YTF_OK;
}

// Activation workspace for `rows` rows (training keeps every layer's Z and H).
struct MlpAct {
  float *part;  // split-K partial sums of the weight gradients
  __bf16 *X0;
  std::vector<float *> Z;
  std::vector<__bf16 *> H;
  float *Y;
  __bf16 *D[2];
};

static int mlp_workspace(ytf_ctx *c, ytf_mlp *m, int64_t rows, bool train, MlpAct &A) {
  std::vector<size_t> sz;
  int maxw = 0;
  for (int l = 0; l < m->nl; ++l) maxw = std::max(maxw, std::max(m->Np[l], m->Kp[l]));
  const size_t part_b = train ? align256(sizeof(float) * 64 * (size_t)maxw * maxw) : 0;
  sz.push_back(align256(sizeof(__bf16) * rows * m->Kp[0]));
  const int nh = m->nl - 1;
  for (int l = 0; l < nh; ++l) sz.push_back(train ? align256(sizeof(float) * rows * m->Np[l]) : 0);
  for (int l = 0; l < nh; ++l) sz.push_back(train || l < 2 ? align256(sizeof(__bf16) * rows * m->Np[l]) : 0);
  sz.push_back(align256(sizeof(float) * rows * m->Np[m->nl - 1]));
  for (int i = 0; i < 2; ++i) sz.push_back(train ? align256(sizeof(__bf16) * rows * maxw) : 0);
  size_t tot = part_b;
  for (size_t b : sz) tot += b;
  int rc = ensure_buf(c, &m->act, &m->act_bytes, tot);
  if (rc) return rc;
  char *p = (char *)m->act + part_b;
  A.part = train ? (float *)m->act : nullptr;
  size_t i = 0;
  A.X0 = (__bf16 *)p;
  p += sz[i++];
  A.Z.assign(nh, nullptr);
  A.H.assign(nh, nullptr);
  for (int l = 0; l < nh; ++l) {
    A.Z[l] = sz[i] ? (float *)p : nullptr;
    p += sz[i++];
  }
  for (int l = 0; l < nh; ++l) {
    A.H[l] = (__bf16 *)p;
    p += sz[i++];
  }
  if (!train) {
    // inference ping-pongs the two hidden buffers that were sized
    char *h0 = (char *)m->act + part_b + sz[0];
    for (int l = 0; l < nh; ++l) A.H[l] = (__bf16 *)(h0 + (l % 2) * sz[1 + nh]);
  }
  A.Y = (float *)p;
  p += sz[i++];
  A.D[0] = (__bf16 *)p;
  p += sz[i++];
  A.D[1] = (__bf16 *)p;
  return YTF_OK;
}

static hipError_t mlp_forward_rows(ytf_ctx *c, ytf_mlp *m, const MlpAct &A, int64_t n, int64_t rows, const float *x,
                                   bool save_z) {
  hipError_t e =
      ytf::launch_mlp_inputs(n, rows, m->nin, m->Kp[0], m->kzf, x, m->norm, m->norm + m->nin, A.X0, c->stream);
  for (int l = 0; l < m->nl && e == hipSuccess; ++l) {
    ytf::MlpGemm g{};
    g.A = l == 0 ? A.X0 : A.H[l - 1];
    g.lda = m->Kp[l];
    g.B = m->P16 + m->offW[l];
    g.ldb = m->Kp[l];
    g.M = (int)rows;
    g.N = m->Np[l];
    g.K = m->Kp[l];
    g.ldc = m->Np[l];
    g.bias = m->P + m->offb[l];
    if (l < m->nl - 1) {
      g.C32 = save_z ? A.Z[l] : nullptr;
      g.C16 = A.H[l];
      e = timed(c, "mlp_gemm", [&] { return ytf::launch_mlp_gemm(g, ytf::MLP_EPI_BIAS_SILU, false, false, c->stream); });
    } else {
      g.C32 = A.Y;
      e = timed(c, "mlp_gemm", [&] { return ytf::launch_mlp_gemm(g, ytf::MLP_EPI_BIAS_F32, false, false, c->stream); });
    }
  }
  return e;
}

// forward + loss + backward into m->G (no update)
static int mlp_loss_grad_locked(ytf_ctx *c, ytf_mlp *m, int64_t n, const float *x, const float *y, const char *who) {
  if (n < 1 || n > ((int64_t)1 << 24) || !x || !y) return fail(c, YTF_EINVAL, "%s: bad batch (n=%lld)", who, (long long)n);
  const int64_t rows = (int64_t)ru64(n);
  MlpAct A;
  int rc = mlp_workspace(c, m, rows, true, A);
  if (rc) return rc;
  hipError_t e = mlp_forward_rows(c, m, A, n, rows, x, true);
  const int L = m->nl - 1;
  const float *osh = m->norm + 2 * m->nin, *osc = osh + m->nout;
  if (e == hipSuccess) e = ytf::launch_mlp_zero(1, m->loss, c->stream);
  if (e == hipSuccess)
    e = ytf::launch_mlp_mse(n, rows, m->nout, m->Np[L], A.Y, y, osh, osc, A.D[L % 2], m->loss, c->stream);
  for (int l = L; l >= 0 && e == hipSuccess; --l) {
    __bf16 *dZ = A.D[l % 2];
    const __bf16 *Hin = l == 0 ? A.X0 : A.H[l - 1];
    // dW[n][k] = sum_b dZ[b][n] Hin[b][k]
    ytf::MlpGemm g{};
    g.A = dZ;
    g.lda = m->Np[l];
    g.B = Hin;
    g.ldb = m->Kp[l];
    g.M = m->Np[l];
    g.N = m->Kp[l];
    g.K = (int)rows;
    g.ldc = m->Kp[l];
    // split the batch reduction so that the small weight tiles fill the chip:
    // partial sums per slice in A.part, then a fixed-order sum into G
    int splits = 1;
    while (splits < 64 && (rows % (64 * splits * 2)) == 0 &&
           (int64_t)(g.M / 64) * (g.N / 64) * splits < 512)
      splits *= 2;
    g.ksplit = splits;
    g.split_stride = (int64_t)g.M * g.N;
    g.C32 = splits > 1 ? A.part : m->G + m->offW[l];
    e = timed(c, "mlp_gemm", [&] { return ytf::launch_mlp_gemm(g, ytf::MLP_EPI_F32, true, true, c->stream); });
    if (e == hipSuccess && splits > 1)
      e = ytf::launch_mlp_reduce(g.split_stride, splits, g.split_stride, A.part, m->G + m->offW[l], c->stream);
    if (e == hipSuccess) e = ytf::launch_mlp_colsum(rows, m->Np[l], dZ, A.part, m->G + m->offb[l], c->stream);
    if (e == hipSuccess && l > 0) {
      // dZ_{l-1} = (dZ_l W_l) * SiLU'(Z_{l-1})
      ytf::MlpGemm h{};
      h.A = dZ;
      h.lda = m->Np[l];
      h.B = m->P16 + m->offW[l];
      h.ldb = m->Kp[l];
      h.M = (int)rows;
      h.N = m->Kp[l];
      h.K = m->Np[l];
      h.ldc = m->Kp[l];
      h.Zin = A.Z[l - 1];
      h.C16 = A.D[(l - 1) % 2];
      e = timed(c, "mlp_gemm", [&] { return ytf::launch_mlp_gemm(h, ytf::MLP_EPI_DSILU, false, true, c->stream); });
    }
  }
  if (e != hipSuccess) return fail(c, YTF_EHIP, "%s: launch failed: %s", who, hipGetErrorString(e));
  return YTF_OK;
}

int ytf_mlp_loss_grad(ytf_ctx *c, ytf_mlp *m, int64_t n, const float *x_dev, const float *y_dev, float *loss_host,
                      float *grad_host) {
  if (!c || !m) return fail(c, YTF_EINVAL, "ytf_mlp_loss_grad: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(c, hipSetDevice(c->device));
  int rc = mlp_loss_grad_locked(c, m, n, x_dev, y_dev, "ytf_mlp_loss_grad");
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (loss_host) HIPCHK(c, hipMemcpy(loss_host, m->loss, sizeof(float), hipMemcpyDeviceToHost));
  if (grad_host) {
    std::vector<float> h(m->np);
    HIPCHK(c, hipMemcpy(h.data(), m->G, sizeof(float) * m->np, hipMemcpyDeviceToHost));
    for (int l = 0; l < m->nl; ++l) {
      for (int i = 0; i < m->N[l]; ++i)
        for (int j = 0; j < m->K[l]; ++j)
          grad_host[m->foffW[l] + (size_t)i * m->K[l] + j] = h[m->offW[l] + (size_t)i * m->Kp[l] + j];
      for (int i = 0; i < m->N[l]; ++i) grad_host[m->foffb[l] + i] = h[m->offb[l] + i];
    }
  }
  return YTF_OK;
}

int ytf_mlp_train_step(ytf_ctx *c, ytf_mlp *m, int64_t n, const float *x_dev, const float *y_dev, float lr,
                       float *loss_dev) {
  if (!c || !m) return fail(c, YTF_EINVAL, "ytf_mlp_train_step: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!(lr > 0.0f) || !std::isfinite(lr)) return fail(c, YTF_EINVAL, "ytf_mlp_train_step: lr must be > 0");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = mlp_loss_grad_locked(c, m, n, x_dev, y_dev, "ytf_mlp_train_step");
  if (rc) return rc;
  m->step += 1;
  m->version += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
  hipError_t e = timed(c, "mlp_adam", [&] {
    return ytf::launch_mlp_adam((int64_t)m->np, m->P, m->Mo, m->Vo, m->G, m->P16, lr, b1, b2, 1e-8f, c1, c2, c->stream);
  });
  if (e == hipSuccess && loss_dev) e = hipMemcpyAsync(loss_dev, m->loss, sizeof(float), hipMemcpyDeviceToDevice, c->stream);
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_mlp_train_step: %s", hipGetErrorString(e));
  return YTF_OK;
}

static ytf::MlpFused mlp_fused_params(const ytf_mlp *m, int64_t n) {
  ytf::MlpFused p{};
  p.n = n;
  p.nin = m->nin;
  p.kzf = m->kzf;
  p.nl = m->nl;
  p.nout = m->nout;
  p.kp0 = m->Kp[0];
  p.W16 = m->P16pk;
  p.P32 = m->P;
  for (int l = 0; l < m->nl; ++l) {
    p.offW[l] = (int64_t)m->offW[l];
    p.offb[l] = (int64_t)m->offb[l];
  }
  p.norm = m->norm;
  return p;
}

static bool mlp_fused_ok(const ytf_mlp *m) {
  return ytf::mlp_fused_supported(m->width, m->nin + 2 * m->kzf, m->nout, m->nl);
}

// Repack the weights into fragment order when they changed since the last pack.
static hipError_t mlp_pack(ytf_ctx *c, ytf_mlp *m) {
  if (m->packed_version == m->version) return hipSuccess;
  hipError_t e = hipSuccess;
  for (int l = 0; l < m->nl && e == hipSuccess; ++l)
    e = ytf::launch_mlp_pack(m->Np[l], m->Kp[l], m->P16 + m->offW[l], m->P16pk + m->offW[l], c->stream);
  if (e == hipSuccess) m->packed_version = m->version;
  return e;
}

int ytf_mlp_forward(ytf_ctx *c, ytf_mlp *m, int64_t n, const float *x_dev, float *y_dev) {
  if (!c || !m) return fail(c, YTF_EINVAL, "ytf_mlp_forward: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n < 0 || (n > 0 && (!x_dev || !y_dev))) return fail(c, YTF_EINVAL, "ytf_mlp_forward: bad arguments");
  if (n == 0) return YTF_OK;
  HIPCHK(c, hipSetDevice(c->device));
  if (mlp_fused_ok(m)) {  // the whole network in one kernel, activations in LDS
    ytf::MlpFused p = mlp_fused_params(m, n);
    p.x = x_dev;
    p.y = y_dev;
    hipError_t e = mlp_pack(c, m);
    if (e == hipSuccess)
      e = timed(c, "mlp_fused", [&] { return ytf::launch_mlp_fused(p, m->width, false, c->stream); });
    if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_mlp_forward: launch failed: %s", hipGetErrorString(e));
    return YTF_OK;
  }
  const int64_t chunk = 1 << 18;
  const int64_t rows_cap = std::min<int64_t>((int64_t)ru64(n), chunk);
  MlpAct A;
  int rc = mlp_workspace(c, m, rows_cap, false, A);
  if (rc) return rc;
  const int L = m->nl - 1;
  const float *osh = m->norm + 2 * m->nin, *osc = osh + m->nout;
  for (int64_t r0 = 0; r0 < n; r0 += chunk) {
    const int64_t nc = std::min(chunk, n - r0), rows = (int64_t)ru64(nc);
    hipError_t e = mlp_forward_rows(c, m, A, nc, rows, x_dev + r0 * m->nin, false);
    if (e == hipSuccess) e = ytf::launch_mlp_outputs(nc, m->nout, m->Np[L], A.Y, osh, osc, y_dev + r0 * m->nout, c->stream);
    if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_mlp_forward: launch failed: %s", hipGetErrorString(e));
  }
  return YTF_OK;
}

int ytf_mlp_set_pole_windows(ytf_ctx *c, ytf_mlp *m, int n_windows, const double *windows) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_mlp_set_pole_windows: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!m || n_windows < 0 || n_windows > 4096 || (n_windows > 0 && !windows))
    return fail(c, YTF_EINVAL, "ytf_mlp_set_pole_windows: bad arguments (at most 4096 windows)");
  for (int t = 0; t < n_windows; ++t)
    if (!(windows[2 * t] <= windows[2 * t + 1]))
      return fail(c, YTF_EINVAL, "ytf_mlp_set_pole_windows: window %d has lo > hi", t);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (m->pole_win) HIPCHK(c, hipFree(m->pole_win));
  m->pole_win = nullptr;
  m->n_pole_win = 0;
  if (n_windows == 0) return YTF_OK;
  HIPCHK(c, hipMalloc((void **)&m->pole_win, sizeof(double) * 2 * (size_t)n_windows));
  HIPCHK(c, hipMemcpy(m->pole_win, windows, sizeof(double) * 2 * (size_t)n_windows, hipMemcpyHostToDevice));
  m->n_pole_win = n_windows;
  return YTF_OK;
}

int ytf_mlp_surrogate_ghat(ytf_ctx *c, ytf_mlp *m, const ytf_layer *layers, int n_layers, double omega,
                           const double *kx_dev, const double *ky_dev, int64_t n, const float *xconst, double zsum,
                           double zsrc, double zobs, double branch_tol, double origin_tol, double t_big, int fallback,
                           double *ghat_dev, unsigned char *flags_dev, int64_t *n_flagged) {
  if (!c || !m) return fail(c, YTF_EINVAL, "ytf_mlp_surrogate_ghat: NULL argument");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n < 0 || n > INT32_MAX || (n > 0 && (!kx_dev || !ky_dev || !ghat_dev || !flags_dev)) ||
      (m->nin > 2 && !xconst))
    return fail(c, YTF_EINVAL, "ytf_mlp_surrogate_ghat: bad arguments");
  if (m->nout != 18 || m->nin < 2 || !mlp_fused_ok(m))
    return fail(c, YTF_EINVAL,
                "ytf_mlp_surrogate_ghat: needs 18 outputs, >= 2 inputs and a hidden width of 128 or 256 (got %d, %d, %d)",
                m->nout, m->nin, m->width);
  if (fallback && (!layers || n_layers < 2)) return fail(c, YTF_EINVAL, "ytf_mlp_surrogate_ghat: no stack for the fallback");
  if (n_flagged) *n_flagged = 0;
  if (n == 0) return YTF_OK;
  HIPCHK(c, hipSetDevice(c->device));
  // The rows flagged from their inputs alone (light line, origin: k_sur_preflag)
  // are known before the network runs: their exact solve (row-list node solve,
  // writing g~ in place) needs no host round trip, and the network leaves their
  // g~ alone.  The rows the network flags (large response) are solved after it,
  // their count also read on the device: nothing waits on the host unless the
  // flagged count is asked for.
  ytf::NodeParams P;
  if (fallback) {  // validates and sends the stack before this call's first launch
    int rc = node_params_locked(c, layers, n_layers, omega, kx_dev, ky_dev, n, 1, 1, 1, &zsrc, &zobs, ghat_dev,
                                nullptr, YTF_TERMS_SCATTERED, 0.0, 1, &P);
    if (rc) return rc;
  }
  // counters (input-flagged, response-flagged) and the two row lists
  const size_t cb = align256(2 * sizeof(int)), lb = align256(sizeof(int) * (size_t)n);
  int rc = ensure_buf(c, &c->aux, &c->aux_bytes, cb + 2 * lb);
  if (rc) return rc;
  int *count = (int *)c->aux;
  int *list_in = (int *)((char *)c->aux + cb), *list_big = (int *)((char *)c->aux + cb + lb);
  ytf::MlpFused p = mlp_fused_params(m, n);
  p.kx = kx_dev;
  p.ky = ky_dev;
  for (int i = 0; i + 2 < m->nin; ++i) p.xconst[i] = xconst[i];
  p.omega = omega;
  p.zsum = zsum;
  p.tbig = t_big;
  p.skip_flagged = fallback ? 1 : 0;
  p.ghat = ghat_dev;
  p.flags = flags_dev;
  p.n_flagged = count + 1;
  p.flagged = list_big;
  auto launch_err = [&](hipError_t e) {
    return fail(c, YTF_EHIP, "ytf_mlp_surrogate_ghat: launch failed: %s", hipGetErrorString(e));
  };
  hipError_t e = ytf::launch_mlp_zero(2, (float *)count, c->stream);  // zero bits = integer 0
  if (e == hipSuccess)
    e = ytf::launch_sur_preflag(n, kx_dev, ky_dev, omega, branch_tol, origin_tol, m->pole_win, m->n_pole_win,
                                flags_dev, count, list_in, c->stream);
  if (e != hipSuccess) return launch_err(e);
  // exact solve of the rows list[0 .. *cnt), g~ written in place (a grid of at
  // most 1024 one-wave blocks; the kernel strides over the count)
  auto solve_rows = [&](const int *list, const int *cnt) -> hipError_t {
    ytf::NodeParams Q = P;
    Q.kx = kx_dev;
    Q.ky = ky_dev;
    Q.nn = n;
    Q.ghat = (cd *)ghat_dev;
    Q.rows = list;
    Q.n_rows = cnt;
    const int blocks = (int)std::min<int64_t>((n + 63) / 64, 1024);
    return timed(c, "node_solve", [&] {
      return (Q.src == 0 && Q.obs == 0) ? ytf::launch_node_top(Q, blocks, c->stream)
                                        : ytf::launch_node_general(Q, blocks, c->stream);
    });
  };
  // (On a second stream beside the network this solve is starved: the
  // network's blocks take every CU slot it frees, and it ends after the network.)
  if (fallback) {
    e = solve_rows(list_in, count);
    if (e != hipSuccess) return launch_err(e);
  }
  e = mlp_pack(c, m);
  if (e == hipSuccess)
    e = timed(c, "mlp_fused", [&] { return ytf::launch_mlp_fused(p, m->width, true, c->stream); });
  if (e != hipSuccess) return launch_err(e);
  if (fallback) {  // the rows the network flagged, counted on the device
    e = solve_rows(list_big, count + 1);
    if (e != hipSuccess) return launch_err(e);
  }
  if (n_flagged) {
    rc = ensure_pinned(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->pinned, count, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *n_flagged = (int64_t)c->pinned[0] + c->pinned[1];
  }
  return YTF_OK;
}

int ytf_ktable_ghat(ytf_ctx *c, const ytf_layer *layers, int n_layers, double omega, int n_panels,
                    const double *kb_dev, const int32_t *kind_dev, const double *sab_dev, const int64_t *row0_dev,
                    const int32_t *m_dev, const double *val_dev, const double *kx_dev, const double *ky_dev,
                    int64_t n, double zsum, double zsrc, double zobs, int fallback, double *ghat_dev,
                    unsigned char *flags_dev, int64_t *n_flagged) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_ktable_ghat: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n < 0 || n > INT32_MAX || (n > 0 && (!kx_dev || !ky_dev || !ghat_dev || !flags_dev)))
    return fail(c, YTF_EINVAL, "ytf_ktable_ghat: bad arguments");
  if (n_panels < 1 || n_panels > 2048 || !kb_dev || !kind_dev || !sab_dev || !row0_dev || !m_dev || !val_dev)
    return fail(c, YTF_EINVAL, "ytf_ktable_ghat: bad table (n_panels=%d must be in 1..2048, no NULL arrays)",
                n_panels);
  if (!(omega > 0)) return fail(c, YTF_EINVAL, "ytf_ktable_ghat: omega must be > 0");
  if (fallback && (!layers || n_layers < 2)) return fail(c, YTF_EINVAL, "ytf_ktable_ghat: no stack for the fallback");
  if (n_flagged) *n_flagged = 0;
  if (n == 0) return YTF_OK;
  HIPCHK(c, hipSetDevice(c->device));
  ytf::NodeParams P;
  if (fallback) {
    int rc = node_params_locked(c, layers, n_layers, omega, kx_dev, ky_dev, n, 1, 1, 1, &zsrc, &zobs, ghat_dev,
                                nullptr, YTF_TERMS_SCATTERED, 0.0, 1, &P);
    if (rc) return rc;
  }
  const size_t cb = align256(sizeof(int)), lb = align256(sizeof(int) * (size_t)n);
  int rc = ensure_buf(c, &c->aux, &c->aux_bytes, cb + lb);
  if (rc) return rc;
  int *count = (int *)c->aux;
  int *list = (int *)((char *)c->aux + cb);
  ytf::KTable T{n_panels, kb_dev, (const int *)kind_dev, sab_dev, row0_dev, (const int *)m_dev, val_dev, 0.0};
  hipError_t e = ytf::launch_mlp_zero(1, (float *)count, c->stream);  // zero bits = integer 0
  if (e == hipSuccess)
    e = timed(c, "ktable", [&] {
      return ytf::launch_ktable_ghat(n, kx_dev, ky_dev, omega, zsum, T, (cd *)ghat_dev, flags_dev, count, list,
                                     c->stream);
    });
  if (e == hipSuccess && fallback) {
    ytf::NodeParams Q = P;
    Q.rows = list;
    Q.n_rows = count;
    const int blocks = (int)std::min<int64_t>((n + 63) / 64, 1024);
    e = timed(c, "node_solve", [&] {
      return (Q.src == 0 && Q.obs == 0) ? ytf::launch_node_top(Q, blocks, c->stream)
                                        : ytf::launch_node_general(Q, blocks, c->stream);
    });
  }
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_ktable_ghat: launch failed: %s", hipGetErrorString(e));
  if (n_flagged) {
    rc = ensure_pinned(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->pinned, count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *n_flagged = (int64_t)c->pinned[0];
  }
  return YTF_OK;
}

// ---------------------------------------------------------------------------
// 2D type-1 nonuniform FFT (nufft.hip)
// ---------------------------------------------------------------------------
// Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n)
static void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(pi * (i + 0.75) / (n + 0.5)), pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double dz = p1 / pp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}

// (-1)^k h / psi(k), k = -N/2 .. N - N/2 - 1, psi(k) = a int phi(z) cos(k a z) dz
static std::vector<double> nu_deconv(int N, int nf, int w, double beta) {
  const double h = 2.0 * 3.14159265358979323846 / nf, a = 0.5 * w * h;
  std::vector<double> zq, wq;
  gauss_legendre(4 * w + 80, zq, wq);
  std::vector<double> phi(zq.size());
  for (size_t q = 0; q < zq.size(); ++q) phi[q] = std::exp(beta * (std::sqrt(1.0 - zq[q] * zq[q]) - 1.0));
  std::vector<double> d(N);
  for (int kk = 0; kk < N; ++kk) {
    const int k = kk - N / 2;
    double s = 0.0;
    for (size_t q = 0; q < zq.size(); ++q) s += wq[q] * phi[q] * std::cos(k * a * zq[q]);
    const double psi = a * s;
    d[kk] = ((k % 2) ? -1.0 : 1.0) * h / psi;
  }
  return d;
}

// The spreading kernel phi(z) = exp(beta (sqrt(1 - z^2) - 1)) in piecewise-
// polynomial form: a point at fine-grid coordinate u with first tap
// mu = ceil(u - w/2) has tap values phi(2 (t + i) / w - 1), i = 0 .. w - 1,
// t = mu - u + w/2 in [0, 1).  Each tap is interpolated in s = 2 t - 1 at
// w + 3 Chebyshev points and stored as monomial coefficients, d-major
// (coef[d w + i], d = 0 .. w + 2).  The largest deviation, 1e-10 at w = 10
// (about e^-beta), sits at the support's ends, where phi has a square-root
// edge; inside it is below 1e-13.
static std::vector<double> nu_horner(int w, double beta) {
  const int n = w + 3;
  const long double pi = 3.141592653589793238462643383279502884L;
  std::vector<double> out((size_t)n * w);
  for (int i = 0; i < w; ++i) {
    std::vector<long double> f(n), cc(n), m(n, 0.0L), t0(n, 0.0L), t1(n, 0.0L), t2(n, 0.0L);
    for (int k = 0; k < n; ++k) {
      const long double x = cosl(pi * (k + 0.5L) / n), z = 2.0L * ((x + 1.0L) / 2.0L + i) / w - 1.0L;
      const long double q = 1.0L - z * z;
      f[k] = q > 0.0L ? expl((long double)beta * (sqrtl(q) - 1.0L)) : 0.0L;
    }
    for (int j = 0; j < n; ++j) {
      long double sum = 0.0L;
      for (int k = 0; k < n; ++k) sum += f[k] * cosl(pi * j * (k + 0.5L) / n);
      cc[j] = sum * 2.0L / n;
    }
    cc[0] /= 2.0L;
    // sum_j cc[j] T_j(s) in monomials: T_0 = 1, T_1 = s, T_{j+1} = 2 s T_j - T_{j-1}
    t0[0] = 1.0L;
    if (n > 1) t1[1] = 1.0L;
    for (int d = 0; d < n; ++d) m[d] = cc[0] * t0[d] + (n > 1 ? cc[1] * t1[d] : 0.0L);
    for (int j = 2; j < n; ++j) {
      for (int d = 0; d < n; ++d) t2[d] = (d ? 2.0L * t1[d - 1] : 0.0L) - t0[d];
      for (int d = 0; d < n; ++d) m[d] += cc[j] * t2[d];
      t0.swap(t1);
      t1.swap(t2);
    }
    for (int d = 0; d < n; ++d) out[(size_t)d * w + i] = (double)m[d];
  }
  return out;
}

int ytf_nufft2d1(ytf_ctx *c, int64_t M, const double *x_dev, const double *y_dev, int ntrans, const double *c_dev,
                 int N1, int N2, double eps, double *f_dev) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_nufft2d1: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  int rc = YTF_OK;
  timed(c, "nufft_apply", [&] {
    rc = nufft2d1_locked(c, M, x_dev, y_dev, ntrans, c_dev, N1, N2, eps, f_dev);
    return hipSuccess;
  });
  return rc;
}

}  // extern "C"

int nufft2d1_locked(ytf_ctx *c, int64_t M, const double *x_dev, const double *y_dev, int ntrans, const double *c_dev,
                    int N1, int N2, double eps, double *f_dev) {
  if (M < 0 || M > ((int64_t)1 << 31) - 1 || ntrans < 1 || N1 < 2 || N2 < 2 || N1 > 2048 || N2 > 2048)
    return fail(c, YTF_EINVAL, "ytf_nufft2d1: bad sizes (M=%lld, ntrans=%d, N1=%d, N2=%d; modes 2..2048)",
                (long long)M, ntrans, N1, N2);
  if (!(eps >= 1e-15 && eps < 1.0)) return fail(c, YTF_EINVAL, "ytf_nufft2d1: eps=%g must be in [1e-15, 1)", eps);
  if (!f_dev || (M > 0 && (!x_dev || !y_dev || !c_dev))) return fail(c, YTF_EINVAL, "ytf_nufft2d1: NULL buffer");
  HIPCHK(c, hipSetDevice(c->device));
  const int w = ytf::nu_width(eps);
  const double beta = 2.30 * w;
  int nf = 64;
  while (nf < 2 * N1 || nf < 2 * N2 || nf < 4 * w) nf *= 2;
  const int nt = nf / 32, ntiles = nt * nt;
  const int64_t Mc = M > 0 ? M : 1;
  const int nsub = ytf::nu_max_subproblems(M, nf);
  // transforms per batch: fine grids and row-pass buffers within ~512 MiB
  const size_t per_t = sizeof(cd) * ((size_t)nf * nf + (size_t)N1 * nf);
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)ntrans, (int64_t)((512u << 20) / per_t), 65535}));
  const size_t uv_b = align256(sizeof(double) * 2 * Mc), i_b = align256(sizeof(int) * Mc),
               t_b = align256(sizeof(int) * (ntiles + 1)), s_b = align256(sizeof(int) * nsub),
               h_b = align256(sizeof(int) * (size_t)std::max(ytf::nu_bin_blocks(Mc), 1) * ntiles),
               g_b = align256(sizeof(cd) * (size_t)nf * nf * nb), T_b = align256(sizeof(cd) * (size_t)N1 * nf * nb);
  int rc = ensure_buf(c, &c->work, &c->work_bytes, uv_b + 3 * i_b + h_b + 3 * t_b + 2 * s_b + g_b + T_b);
  if (rc) return rc;
  char *p = (char *)c->work;
  double *uv = (double *)p; p += uv_b;
  int *tile = (int *)p; p += i_b;
  int *rank = (int *)p; p += i_b;
  int *order = (int *)p; p += i_b;
  int *H = (int *)p; p += h_b;
  int *count = (int *)p; p += t_b;
  int *start = (int *)p; p += t_b;
  int *substart = (int *)p; p += t_b;
  int *sub_tile = (int *)p; p += s_b;
  int *sub_p0 = (int *)p; p += s_b;
  cd *grid = (cd *)p; p += g_b;
  cd *T1 = (cd *)p;
  double *dd1 = nullptr, *dd2 = nullptr;
  for (int dim = 0; dim < 2; ++dim) {
    const int N = dim ? N2 : N1;
    auto key = std::make_tuple(N, nf, w);
    auto it = c->nu_deconv.find(key);
    if (it == c->nu_deconv.end()) {
      const std::vector<double> d = nu_deconv(N, nf, w, beta);
      double *dp = nullptr;
      HIPCHK(c, hipMalloc(&dp, sizeof(double) * N));
      HIPCHK(c, hipMemcpy(dp, d.data(), sizeof(double) * N, hipMemcpyHostToDevice));
      it = c->nu_deconv.emplace(key, dp).first;
    }
    (dim ? dd2 : dd1) = it->second;
  }
  auto hit = c->nu_horner.find(w);
  if (hit == c->nu_horner.end()) {
    const std::vector<double> hv = nu_horner(w, beta);
    double *hp = nullptr;
    HIPCHK(c, hipMalloc(&hp, sizeof(double) * hv.size()));
    HIPCHK(c, hipMemcpy(hp, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice));
    hit = c->nu_horner.emplace(w, hp).first;
  }
  const double *hc = hit->second;
  cd *tw = nullptr;
  rc = twiddles(c, nf, &tw);
  if (rc) return rc;
  hipError_t e = timed(c, "nufft_sort", [&] {
    return ytf::launch_nufft_sort(M, x_dev, y_dev, nf, uv, tile, rank, H, count, start, substart, sub_tile, sub_p0,
                                  order, c->stream);
  });
  for (int t0 = 0; t0 < ntrans && e == hipSuccess; t0 += nb) {
    const int b = std::min(nb, ntrans - t0);
    e = timed(c, "nufft_spread", [&] {
      return ytf::launch_nufft_spread(M, b, nf, w, hc, start, substart, sub_tile, sub_p0, order, uv,
                                      (const cd *)c_dev + (size_t)t0 * Mc, Mc, grid, c->stream);
    });
    if (e == hipSuccess)
      e = timed(c, "nufft_fft", [&] {
        return ytf::launch_nufft_fft(b, nf, grid, T1, (cd *)f_dev + (size_t)t0 * N1 * N2, N1, N2, dd1, dd2, tw,
                                     c->stream);
      });
  }
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_nufft2d1: launch failed: %s", hipGetErrorString(e));
  return YTF_OK;
}

extern "C" {

static int radial_synthesis_locked(ytf_ctx *c, const char *who, int64_t n, const double *kx, const double *wt,
                                   int np, const double *ghat, int nr, const double *rho, cd *R, int mode) {
  if (mode != YTF_SYNTH_FP64 && mode != YTF_SYNTH_F16X3)
    return fail(c, YTF_EINVAL, "%s: unknown precision mode %d", who, mode);
  if (n < 0 || np < 1 || nr < 0) return fail(c, YTF_EINVAL, "%s: bad sizes (n_nodes=%lld, n_pairs=%d, n_rho=%d)", who,
                                             (long long)n, np, nr);
  if (nr == 0) return YTF_OK;
  if (n > 0 && (!kx || !wt || !ghat)) return fail(c, YTF_EINVAL, "%s: NULL node buffer", who);
  const size_t coef_b = align256(sizeof(cd) * 9 * (size_t)np * (size_t)(n > 0 ? n : 1));
  hipError_t e = hipSuccess;
  if (mode == YTF_SYNTH_F16X3 && n > 0) {
    const ytf::MfmaPlan P = ytf::mfma_plan(n, np, nr);
    const size_t jb = align256(P.jtab_bytes), cb = align256(P.ctab_bytes), eb = align256(P.cexp_bytes),
                 ob = align256(P.out_bytes), sb = align256(P.split_bytes);
    int rc = ensure_buf(c, &c->work, &c->work_bytes, coef_b + jb + cb + eb + ob + sb);
    if (rc) return rc;
    char *w = (char *)c->work;
    cd *coef = (cd *)w;
    void *jtab = w + coef_b, *ctab = w + coef_b + jb;
    int *cexp = (int *)(w + coef_b + jb + cb);
    float *out = (float *)(w + coef_b + jb + cb + eb);
    int *split = (int *)(w + coef_b + jb + cb + eb + ob);
    if ((e = timed(c, "synth_coef", [&] { return ytf::launch_synth_coef(n, np, kx, wt, (const cd *)ghat, coef, c->stream); })) ||
        (e = timed(c, "bessel_table", [&] { return ytf::launch_bessel_table(n, np, nr, kx, rho, jtab, c->stream); })) ||
        (e = timed(c, "split_columns", [&] { return ytf::launch_split_columns(n, np, nr, coef, ctab, cexp, c->stream); })) ||
        (e = timed(c, "bessel_gemm", [&] {
           return ytf::launch_bessel_gemm(n, np, nr, jtab, ctab, out, split, c->n_cu, c->stream);
         })) ||
        (e = timed(c, "mfma_combine", [&] { return ytf::launch_mfma_combine(n, np, nr, out, split, cexp, R, c->stream); })))
      return fail(c, YTF_EHIP, "%s: matrix-core launch failed: %s", who, hipGetErrorString(e));
    return YTF_OK;
  }
  const ytf::SynthPlan SP = ytf::synth_fp64_plan(n, np, nr);
  const size_t jtab_b = align256(sizeof(double) * 3 * (size_t)std::max<int64_t>(SP.batch_nodes, 1) * (size_t)nr),
               part_b = align256(sizeof(cd) * 9 * (size_t)np * (size_t)nr * (size_t)std::max(SP.nchunks, 1)),
               sum_b = align256(sizeof(cd) * 9 * (size_t)np * (size_t)nr);
  int rc = ensure_buf(c, &c->work, &c->work_bytes, coef_b + jtab_b + part_b + sum_b);
  if (rc) return rc;
  char *w = (char *)c->work;
  cd *coef = (cd *)w;
  double *jtab = (double *)(w + coef_b);
  cd *part = (cd *)(w + coef_b + jtab_b);
  cd *S = (cd *)(w + coef_b + jtab_b + part_b);
  if (n > 0 &&
      (e = timed(c, "synth_coef", [&] { return ytf::launch_synth_coef(n, np, kx, wt, (const cd *)ghat, coef, c->stream); })))
    return fail(c, YTF_EHIP, "%s: coefficient launch failed: %s", who, hipGetErrorString(e));
  // with no nodes the chunk sum writes zeros
  if ((e = timed(c, "synth_fp64", [&] {
         return ytf::launch_synth_fp64(n, np, nr, SP, kx, rho, coef, jtab, part, S, c->stream);
       })) ||
      (e = timed(c, "synth_reduce", [&] { return ytf::launch_synth_reduce(np, nr, 1, S, R, c->stream); })))
    return fail(c, YTF_EHIP, "%s: synthesis launch failed: %s", who, hipGetErrorString(e));
  return YTF_OK;
}

int ytf_radial_synthesis(ytf_ctx *c, int64_t n_nodes, const double *kx_dev, const double *wt_dev, int n_pairs,
                         const double *ghat_dev, int n_rho, const double *rho_dev, double *R_dev, int mode) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_radial_synthesis: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n_rho > 0 && (!rho_dev || !R_dev)) return fail(c, YTF_EINVAL, "ytf_radial_synthesis: NULL rho/R");
  HIPCHK(c, hipSetDevice(c->device));
  return radial_synthesis_locked(c, "ytf_radial_synthesis", n_nodes, kx_dev, wt_dev, n_pairs, ghat_dev, n_rho,
                                 rho_dev, (cd *)R_dev, mode);
}

int ytf_radial_terms(ytf_ctx *c, int64_t n_nodes, const double *kx_dev, const double *wt_dev, int n_pairs,
                     const double *ghat_dev, int n_rho, const double *rho_dev, double *T_dev) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_radial_terms: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n_nodes < 0 || n_pairs < 0 || n_rho < 0)
    return fail(c, YTF_EINVAL, "ytf_radial_terms: negative size (n_nodes=%lld, n_pairs=%d, n_rho=%d)",
                (long long)n_nodes, n_pairs, n_rho);
  if (n_nodes == 0 || n_pairs == 0 || n_rho == 0) return YTF_OK;
  if (!kx_dev || !wt_dev || !ghat_dev || !rho_dev || !T_dev)
    return fail(c, YTF_EINVAL, "ytf_radial_terms: NULL pointer");
  if (n_nodes * n_rho > ((int64_t)1 << 31) * 255)
    return fail(c, YTF_EINVAL, "ytf_radial_terms: n_nodes x n_rho too large");
  HIPCHK(c, hipSetDevice(c->device));
  hipError_t e = timed(c, "radial_terms", [&] {
    return ytf::launch_synth_terms(n_nodes, n_pairs, n_rho, kx_dev, wt_dev, (const cd *)ghat_dev, rho_dev, (cd *)T_dev,
                                   c->stream);
  });
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_radial_terms: %s", hipGetErrorString(e));
  return YTF_OK;
}

int ytf_closed_form_G_1d(ytf_ctx *c, int64_t n_nodes, const double *kx_dev, const double *wt_dev, int n_pairs,
                         const double *ghat_dev, const double *xs, int nx, const double *ys, int ny, double *G_dev,
                         int mode) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_closed_form_G_1d: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (nx < 0 || ny < 0 || (nx > 0 && !xs) || (ny > 0 && !ys))
    return fail(c, YTF_EINVAL, "ytf_closed_form_G_1d: bad pixel axes (nx=%d, ny=%d)", nx, ny);
  const int64_t npix = (int64_t)nx * ny;
  if (npix == 0) return YTF_OK;
  if (!G_dev) return fail(c, YTF_EINVAL, "ytf_closed_form_G_1d: G is NULL");
  if (npix > (int64_t)1 << 31) return fail(c, YTF_EINVAL, "ytf_closed_form_G_1d: too many pixels");
  HIPCHK(c, hipSetDevice(c->device));
  // distinct radii in first-appearance order over the column-major pixel grid
  std::vector<double> rho;
  std::vector<int> rmap((size_t)npix);
  std::map<double, int> seen;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double r = std::hypot(xs[i], ys[j]);
      auto it = seen.find(r);
      int idx;
      if (it == seen.end()) {
        idx = (int)rho.size();
        seen.emplace(r, idx);
        rho.push_back(r);
      } else {
        idx = it->second;
      }
      rmap[i + (size_t)nx * j] = idx;
    }
  const int nr = (int)rho.size();
  const size_t rb = align256(sizeof(double) * nr), xb = align256(sizeof(double) * nx),
               yb = align256(sizeof(double) * ny), mb = align256(sizeof(int) * (size_t)npix),
               Rb = align256(sizeof(cd) * 9 * (size_t)n_pairs * nr);
  // the previous call's kernels may still read aux: wait before refilling it
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int rc = ensure_buf(c, &c->aux, &c->aux_bytes, rb + xb + yb + mb + Rb);
  if (rc) return rc;
  char *a = (char *)c->aux;
  double *rho_d = (double *)a, *xs_d = (double *)(a + rb), *ys_d = (double *)(a + rb + xb);
  int *rmap_d = (int *)(a + rb + xb + yb);
  cd *R = (cd *)(a + rb + xb + yb + mb);
  HIPCHK(c, hipMemcpyAsync(rho_d, rho.data(), sizeof(double) * nr, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(xs_d, xs, sizeof(double) * nx, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(ys_d, ys, sizeof(double) * ny, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(rmap_d, rmap.data(), sizeof(int) * npix, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  rc = radial_synthesis_locked(c, "ytf_closed_form_G_1d", n_nodes, kx_dev, wt_dev, n_pairs, ghat_dev, nr, rho_d, R,
                               mode);
  if (rc) return rc;
  hipError_t e = timed(c, "synth_rotate", [&] {
    return ytf::launch_synth_rotate(nx, ny, n_pairs, nr, xs_d, ys_d, rmap_d, R, (cd *)G_dev, c->stream);
  });
  if (e != hipSuccess) return fail(c, YTF_EHIP, "ytf_closed_form_G_1d: rotation launch failed: %s", hipGetErrorString(e));
  return YTF_OK;
}

}  // extern "C"
