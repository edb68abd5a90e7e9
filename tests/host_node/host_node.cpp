// Host build of the Green k-node solve for the sanitizer test
// (tests/test_host_node_sanitized.py): node_core.hpp compiled as host C++
// (tests/host_node/include stands in for the HIP header), one lane per node,
// with AddressSanitizer and UndefinedBehaviorSanitizer.  The LDS block of the
// kernel is a heap array that ends right after the last slot the case
// declares (LdsSlots), so a slot index past it is a heap overflow; with
// `poison`, every slot is set to NaN before each node, so a slot read before
// the node wrote it turns the node's g~ into NaN.
//
// Input file (raw, native endianness): int32 nl; nl x ytf_layer; double
// omega; int64 n; n kx; n ky; int32 src, obs (1-based), npairs; npairs zsrc;
// npairs zobs; int32 terms; double dz; int32 ez_local, poison.
// Output file: g~ [npairs][9][n] complex (interleaved doubles).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <limits>
#include <vector>

#include "node_core.hpp"
#include "ytf.h"

using namespace ytf;
static_assert(sizeof(ytf_layer) == sizeof(LayerDev), "ytf_layer and LayerDev share one layout");

// As node_kernel and node_kernel_redo: the fast solve, then the robust solve
// (the interface QR fallback) of a node whose fast solve met an exactly
// singular interface.  Prints the number of such nodes.  The fast solve takes
// the depth pairs one at a time, as node_kernel does with grid.y
// groups (each group solves the node again; the first writes the flags and
// the cache outputs), so a pair-range mistake changes g~.
template <int CASE, bool GEN>
static void run(const NodeParams &P, bool poison) {
  // slot i of lane 0 is element 64 i: the array ends at the last declared slot
  const size_t n_el = (size_t)(LdsSlots<CASE>::n - 1) * 64 + 1;
  std::vector<cd> lds(n_el);
  const LaneMat st{lds.data(), 0};
  const double nan = std::numeric_limits<double>::quiet_NaN();
  int64_t n_redo = 0;
  for (int64_t k = 0; k < P.nn; ++k) {
    if (poison)
      for (auto &v : lds) v = cd{nan, nan};
    bool redo = false;
    for (int p0 = 0; p0 < std::max(P.npairs, 1); ++p0) {
      if (poison && p0 > 0)
        for (auto &v : lds) v = cd{nan, nan};
      bool r = false;
      solve_node<CASE, GEN, false>(P, k, st, r, p0, std::min(P.npairs, p0 + 1));
      if (p0 == 0) redo = r;
    }
    if (redo) {
      ++n_redo;
      if (poison)
        for (auto &v : lds) v = cd{nan, nan};
      bool again = false;
      solve_node<CASE, GEN, true>(P, k, st, again, 0, P.npairs);
    }
  }
  printf("redo %lld\n", (long long)n_redo);
}

template <bool GEN>
static void dispatch(const NodeParams &P, bool poison) {
  if (P.src == 0 && P.obs == 0) run<CASE_TOP, GEN>(P, poison);
  else if (P.obs == P.src) run<CASE_SAME, GEN>(P, poison);
  else if (P.obs > P.src) run<CASE_RIGHT, GEN>(P, poison);
  else run<CASE_LEFT, GEN>(P, poison);
}

template <class T>
static bool rd(FILE *f, T *v, size_t n = 1) {
  return fread(v, sizeof(T), n, f) == n;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: host_node input output\n");
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t nl, src, obs, npairs, terms, ez_local, poison;
  double omega, dz;
  int64_t n;
  bool ok = rd(f, &nl) && nl >= 2;
  std::vector<ytf_layer> layers(ok ? nl : 0);
  ok = ok && rd(f, layers.data(), layers.size()) && rd(f, &omega) && rd(f, &n) && n >= 0;
  std::vector<double> kx(ok ? n : 0), ky(ok ? n : 0);
  ok = ok && rd(f, kx.data(), kx.size()) && rd(f, ky.data(), ky.size()) && rd(f, &src) && rd(f, &obs) &&
       rd(f, &npairs) && npairs >= 1;
  std::vector<double> zs(ok ? npairs : 0), zo(ok ? npairs : 0);
  ok = ok && rd(f, zs.data(), zs.size()) && rd(f, zo.data(), zo.size()) && rd(f, &terms) && rd(f, &dz) &&
       rd(f, &ez_local) && rd(f, &poison);
  fclose(f);
  if (!ok || src < 1 || src >= nl || obs < 1 || obs > nl) {
    fprintf(stderr, "host_node: bad input\n");
    return 2;
  }
  std::vector<cd> g((size_t)npairs * 9 * n);
  NodeParams P;
  P.L = (const LayerDev *)layers.data();
  P.nl = nl;
  P.omega = omega;
  P.kx = kx.data();
  P.ky = ky.data();
  P.nn = n;
  P.src = src - 1;
  P.obs = obs - 1;
  P.npairs = npairs;
  P.zsrc = zs.data();
  P.zobs = zo.data();
  P.ghat = g.data();
  P.Qsrc = P.sf = P.sb = P.T = P.basis = nullptr;
  P.flags = nullptr;
  P.terms = terms;
  P.dz = dz;
  P.ez_local = ez_local;
  // the general 4x4 eigen-solver: a layer with xz / yz coupling (as the C-ABI decides)
  bool general = false;
  for (const ytf_layer &Ly : layers) {
    if ((Ly.eps_kind == 0 || Ly.eps_kind == 1) && Ly.mu_kind == 0) continue;
    for (int e : {2, 5, 6, 7})
      general = general || Ly.eps[2 * e] != 0.0 || Ly.eps[2 * e + 1] != 0.0 || Ly.mu[2 * e] != 0.0 ||
                Ly.mu[2 * e + 1] != 0.0;
  }
  P.needs_general = general;
  if (general) dispatch<true>(P, poison != 0);
  else dispatch<false>(P, poison != 0);
  FILE *o = fopen(argv[2], "wb");
  if (!o || fwrite(g.data(), sizeof(cd), g.size(), o) != g.size()) return 3;
  fclose(o);
  return 0;
}
