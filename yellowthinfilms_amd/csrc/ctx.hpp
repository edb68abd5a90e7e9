// ctx.hpp -- the context behind the C-ABI (one device + one stream) and the
// host-side helpers shared by the C-ABI translation units: error reporting,
// context-owned device buffers, per-launch kernel timing, FFT twiddle tables.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "ytf.h"
#include "ytf_kernels.h"

using ytf::cd;
using ytf::LayerDev;

struct ytf_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
  std::mutex mu;
  // small per-call inputs (layers, z pairs) in a few device slots keyed by
  // their bytes: inputs that alternate between calls (one set per
  // observation layer, say) are sent once, and a call waits for earlier
  // kernels only when it has to overwrite a slot they may still read
  struct SmallSlot {
    std::vector<char> host;  // what the device copy holds (and the copy's source)
    void *dev = nullptr;
    size_t cap = 0;
    uint64_t used = 0;
  };
  static constexpr int kSmallSlots = 8;
  SmallSlot small_slots[kSmallSlots];
  uint64_t small_tick = 0;
  void *small = nullptr;  // the slot of the current call
  // FFT twiddle tables, one per length
  std::map<int, cd *> tw;
  // NUFFT deconvolution factors per (modes, fine grid, kernel width)
  std::map<std::tuple<int, int, int>, double *> nu_deconv;
  // NUFFT spreading kernel's piecewise-polynomial coefficients per width
  std::map<int, double *> nu_horner;
  // scratch of the synthesis kernels and its host-side inputs (radii, pixel
  // coordinates, radius map), grown on demand, freed with the context
  void *work = nullptr;
  size_t work_bytes = 0;
  void *aux = nullptr;
  size_t aux_bytes = 0;
  // redo marks of the node solves (two words, a byte per node): zero between
  // launches, grown on demand
  void *redo = nullptr;
  size_t redo_bytes = 0;
  // pinned host words for counts read back
  int *pinned = nullptr;
  // kernel timing: events bracketing each launch while enabled
  int n_cu = 0;  // compute units of the device
  bool timing = false;
  std::vector<std::string> timing_only;  // non-empty: only launches of these names are bracketed
  struct Mark {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long>> totals;
};

inline thread_local std::string g_err;

inline int fail(ytf_ctx *c, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
inline int fail(ytf_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  g_err = buf;
  return code;
}
#define HIPCHK(c, expr)                                                                           \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) return fail((c), YTF_EHIP, "%s: %s", #expr, hipGetErrorString(_e));     \
  } while (0)

// Point c->small at a device copy of hb: a slot that already holds these
// bytes, else an empty slot, else the least recently used one (after the
// device has drained, since queued kernels, on this stream or on one the
// context was bound to before, may still read it).
inline int set_small(ytf_ctx *c, const std::vector<char> &hb) {
  const uint64_t tick = ++c->small_tick;
  for (auto &s : c->small_slots)
    if (s.dev && s.host == hb) {
      s.used = tick;
      c->small = s.dev;
      return YTF_OK;
    }
  ytf_ctx::SmallSlot *v = nullptr;
  for (auto &s : c->small_slots)
    if (!s.dev) {
      v = &s;
      break;
    }
  if (!v) {
    v = &c->small_slots[0];
    for (auto &s : c->small_slots)
      if (s.used < v->used) v = &s;
    HIPCHK(c, hipDeviceSynchronize());
  }
  if (v->cap < hb.size()) {
    if (v->dev) HIPCHK(c, hipFree(v->dev));  // hipFree waits for the device
    v->dev = nullptr;
    v->cap = 0;
    const size_t nb = hb.size() < 4096 ? 4096 : hb.size();
    HIPCHK(c, hipMalloc(&v->dev, nb));
    v->cap = nb;
  }
  v->host = hb;  // the copy's source lives until the slot is reused, after a drain
  v->used = tick;
  HIPCHK(c, hipMemcpyAsync(v->dev, v->host.data(), hb.size(), hipMemcpyHostToDevice, c->stream));
  c->small = v->dev;
  return YTF_OK;
}

// Grow a context-owned device buffer to at least `bytes` (contents are not kept).
inline int ensure_buf(ytf_ctx *c, void **buf, size_t *cur, size_t bytes) {
  if (bytes <= *cur) return YTF_OK;
  if (*buf) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(*buf));
    *buf = nullptr;
    *cur = 0;
  }
  size_t nb = bytes < (1u << 20) ? (1u << 20) : bytes;
  HIPCHK(c, hipMalloc(buf, nb));
  *cur = nb;
  return YTF_OK;
}

// The node solves' redo marks for n rows (two words and a byte per row;
// zeroed when it grows; each redo launch leaves it zero again).
inline int ensure_redo(ytf_ctx *c, int64_t n) {
  const size_t bytes = 2 * sizeof(int) + (size_t)n;
  if (bytes <= c->redo_bytes) return YTF_OK;
  int rc = ensure_buf(c, &c->redo, &c->redo_bytes, bytes);
  if (rc) return rc;
  HIPCHK(c, hipMemsetAsync(c->redo, 0, c->redo_bytes, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return YTF_OK;
}

// Allocate the pinned words on first use.
inline int ensure_pinned(ytf_ctx *c) {
  if (!c->pinned) HIPCHK(c, hipHostMalloc((void **)&c->pinned, 256, hipHostMallocDefault));
  return YTF_OK;
}

inline hipEvent_t take_event(ytf_ctx *c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  // timing-only events: no system-scope fence (cache write-back and
  // invalidate) when they are recorded, so bracketing a launch costs the
  // launches around it next to nothing; drain_marks synchronizes the stream
  // before it reads them
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
  return e;
}

// Launch through `f`; while timing is on, bracket it with events on stream
// s (timed: the context's stream) under `name`.
template <class F>
inline hipError_t timed_on(ytf_ctx *c, hipStream_t s, const char *name, F &&f) {
  if (!c->timing) return f();
  if (!c->timing_only.empty()) {
    bool hit = false;
    for (const auto &t : c->timing_only) hit = hit || t == name;
    if (!hit) return f();
  }
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return f();
  hipEventRecord(a, s);
  hipError_t e = f();
  hipEventRecord(b, s);
  c->marks.push_back({name, a, b});
  return e;
}
template <class F>
inline hipError_t timed(ytf_ctx *c, const char *name, F &&f) {
  return timed_on(c, c->stream, name, static_cast<F &&>(f));
}

// Fold finished marks into the per-name totals (synchronizes the stream).
inline int drain_marks(ytf_ctx *c) {
  if (c->marks.empty()) return YTF_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &m : c->marks) {
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, m.a, m.b));
    auto &t = c->totals[m.name];
    t.first += ms;
    t.second += 1;
    c->pool.push_back(m.a);
    c->pool.push_back(m.b);
  }
  c->marks.clear();
  return YTF_OK;
}

// The length-n root table exp(-2 pi i j / n), j < n, rounded from long double.
inline void twiddle_table(int n, cd *h) {
  for (int j = 0; j < n; ++j) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
    h[j].r = (double)cosl(a);
    h[j].i = (double)sinl(a);
  }
}

inline int twiddles(ytf_ctx *c, int n, cd **out) {
  auto it = c->tw.find(n);
  if (it != c->tw.end()) { *out = it->second; return YTF_OK; }
  std::vector<cd> h(n);
  twiddle_table(n, h.data());
  cd *d = nullptr;
  HIPCHK(c, hipMalloc(&d, sizeof(cd) * n));
  HIPCHK(c, hipMemcpy(d, h.data(), sizeof(cd) * n, hipMemcpyHostToDevice));
  c->tw[n] = d;
  *out = d;
  return YTF_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Per-device set-up of the allocation-free entry points (noalloc.hip); run by
// ytf_ctx_create with the device current.
int noalloc_init(int device, std::string *err);

// Entry points for callers that already hold the context's lock (capi.hip).
int green_nodes_locked(ytf_ctx *c, const ytf_layer *layers, int n_layers, double omega, const double *kx_dev,
                       const double *ky_dev, int64_t n_nodes, int source_layer, int obs_layer, int n_pairs,
                       const double *zsrc, const double *zobs, double *ghat_dev, const ytf_node_cache *cache,
                       int terms, double dz_box, int ez_local);
// Uses the context's `work` buffer (may reallocate it).
int nufft2d1_locked(ytf_ctx *c, int64_t M, const double *x_dev, const double *y_dev, int ntrans, const double *c_dev,
                    int N1, int N2, double eps, double *f_dev);
