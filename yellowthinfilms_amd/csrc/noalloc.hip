// noalloc.hip -- ytf_spectral_apply_noalloc: the uniform-grid spectral apply of
// ytf_spectral_apply with its work array and root tables in one file-scope
// device array instead of memory the caller or the library allocates, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops).  Every call on a device shares that device's array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"

namespace {

constexpr int kMaxPoints = 1 << 18;  // nx * ny: 512 x 512
constexpr int kMinLen = 8, kMaxLen = 2048;
constexpr size_t kWorkBytes = sizeof(cd) * 3 * (size_t)kMaxPoints;  // three planes of the grid, 12 MiB
constexpr size_t kTwEntries = 2 * kMaxLen - kMinLen;                 // root tables of 8, 16, ..., 2048
constexpr size_t kScratchBytes = kWorkBytes + ((sizeof(cd) * kTwEntries + 255) & ~(size_t)255);

// offset (entries) of the length-n table: the tables of 8 .. n/2 come first
size_t tw_offset(int n) { return (size_t)n - kMinLen; }

bool pow2_len(int n) { return n >= kMinLen && n <= kMaxLen && (n & (n - 1)) == 0; }

std::mutex g_mu;
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

}  // namespace

// The work array (first kWorkBytes) and the root tables.  In host code this
// name only stands in for the symbol: its device address comes from
// hipGetSymbolAddress, per device, in noalloc_init.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

int noalloc_init(int device, std::string *err) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (device < 0 || device >= kMaxDevices) {
    *err = "noalloc set-up: device index out of range";
    return YTF_EINVAL;
  }
  if (g_scratch[device]) return YTF_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    *err = std::string("noalloc set-up: hipGetSymbolAddress: ") + hipGetErrorString(e);
    return YTF_EHIP;
  }
  // root tables, computed as ytf_spectral_apply's are, so both entry points
  // give the same bits
  std::vector<cd> h(kTwEntries);
  for (int n = kMinLen; n <= kMaxLen; n *= 2) twiddle_table(n, h.data() + tw_offset(n));
  e = hipMemcpy((char *)p + kWorkBytes, h.data(), sizeof(cd) * kTwEntries, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    *err = std::string("noalloc set-up: root tables: ") + hipGetErrorString(e);
    return YTF_EHIP;
  }
  g_scratch[device] = p;
  return YTF_OK;
}

extern "C" int ytf_spectral_apply_noalloc(ytf_ctx *c, int nx, int ny, const double *ghat_dev, const double *J_dev,
                                          double *E_dev, double scale) {
  if (!c) return fail(nullptr, YTF_EINVAL, "ytf_spectral_apply_noalloc: ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!pow2_len(nx) || !pow2_len(ny))
    return fail(c, YTF_EINVAL, "ytf_spectral_apply_noalloc: nx=%d, ny=%d unsupported (powers of two %d..%d)", nx, ny,
                kMinLen, kMaxLen);
  if ((int64_t)nx * ny > kMaxPoints)
    return fail(c, YTF_EINVAL,
                "ytf_spectral_apply_noalloc: nx*ny = %lld exceeds the %d points of its fixed work array "
                "(use ytf_spectral_apply)",
                (long long)nx * ny, kMaxPoints);
  if (!ghat_dev || !J_dev || !E_dev) return fail(c, YTF_EINVAL, "ytf_spectral_apply_noalloc: NULL buffer");
  if ((const void *)E_dev == (const void *)J_dev) return fail(c, YTF_EINVAL, "ytf_spectral_apply_noalloc: E aliases J");
  void *p = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
  if (!p) return fail(c, YTF_EHIP, "ytf_spectral_apply_noalloc: device %d not set up", c->device);
  HIPCHK(c, hipSetDevice(c->device));
  cd *work = (cd *)p;
  const cd *tw = (const cd *)((char *)p + kWorkBytes);
  hipError_t e = ytf::launch_spectral_apply(nx, ny, (const cd *)ghat_dev, (const cd *)J_dev, (cd *)E_dev, work,
                                            tw + tw_offset(nx), tw + tw_offset(ny), scale, c->stream);
  if (e != hipSuccess)
    return fail(c, YTF_EHIP, "ytf_spectral_apply_noalloc: launch failed: %s", hipGetErrorString(e));
  return YTF_OK;
}
