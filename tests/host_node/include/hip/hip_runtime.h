// Host stand-ins for the HIP qualifiers and the two wave intrinsics that the
// Green k-node solve (yellowthinfilms_amd/csrc/node_core.hpp) uses, so that
// the same source builds as plain host C++ for the sanitizer test
// (tests/test_host_node_sanitized.py).  Test infrastructure only: it models
// one lane, and nothing built with it runs on a GPU.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__

struct ytf_host_dim3 {
  unsigned x, y, z;
};
static const ytf_host_dim3 threadIdx{0, 0, 0}, blockIdx{0, 0, 0}, blockDim{1, 1, 1}, gridDim{1, 1, 1};
typedef int hipError_t;
typedef void *hipStream_t;
enum { hipSuccess = 0 };

// A wave of one lane: a test over the wave is the lane's own value, and the
// first lane is the lane.
inline bool __all(bool x) { return x; }
inline bool __any(bool x) { return x; }
// one lane, one block: barriers and fences are no-ops, atomics plain updates
inline void __syncthreads() {}
inline void __threadfence() {}
inline int atomicAdd(int *p, int v) {
  const int o = *p;
  *p = o + v;
  return o;
}
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __builtin_amdgcn_rcp(x) (1.0 / (x))

using std::isfinite;
using std::isnan;
using std::max;
using std::min;
