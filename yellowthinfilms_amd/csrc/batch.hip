// batch.hip -- uneven batches of independent stacks and k-nodes in one
// persistent launch, FP64.
//
// Per node (one lane): the Berreman modes of every layer with continuity
// matching (Modes_Stack, mod_ModeSolver.jl:358-413), the interface S-matrices
// (_S_Interface_Matrix, mod_ScatteringMatrix.jl:27-47) and the Redheffer
// cascade to the stack's total S = SLeft[N] = Sprop[1] * Sint[1] * ... *
// Sint[N-1] (S_all_Matrix, :112-154, "*" the star product), plus the modes of
// the first and last layers (the incident and exit media).
//
// Scheduling: work items are runs of `nodes_per_item` consecutive nodes of one
// stack, ordered heaviest stack first (more layers = more work).  The grid is
// persistent: every wave takes items from one atomic counter until none are
// left, so stacks of 2..24 layers and 1k..16k nodes share the chip without a
// tail of idle CUs.  All lanes of a wave work on the same stack, so the layer
// loop is uniform and the layer table is read through scalar loads.
#include <hip/hip_runtime.h>

#include "node_core.hpp"

namespace ytf {

struct BatchArgs {
  const LayerDev *L;        // all layers, stacks concatenated
  const int *stack_l0;      // first layer of stack s
  const int *stack_nl;      // layer count of stack s
  const int64_t *stack_n0;  // first node of stack s
  const int64_t *stack_nn;  // node count of stack s
  const int *item_stack;    // work item -> stack
  const int *item_first;    // work item -> first node within the stack
  int n_items;
  int nodes_per_item;  // 64 * nodes per lane
  double omega;
  const double *kx, *ky;
  cd *S, *V1, *Q1, *VN, *QN;  // [e][ntot] (node fastest); any may be null except S
  unsigned char *flags;       // [ntot]: 1 valid, 4 singular interface (no redo yet), 8 mode failure
  int64_t ntot;
  unsigned int *counter;
  int *redo;  // redo marks (two words, a byte per node)
};

// One node of one stack.  ROBUST = false: an interface LU that meets an
// exactly zero pivot only sets `redo`; the caller solves the node again with
// ROBUST = true, which takes the reference's pivoted-QR fallback there
// (mod_ScatteringMatrix.jl:35-44).
template <bool GENERAL, bool ROBUST>
__device__ void stack_node(const BatchArgs &A, const LayerDev *L, int nl, int64_t g, const LaneMat &Vp, bool &redo) {
  const double kx = A.kx[g], ky = A.ky[g], w = A.omega;
  cd p[4], acc[16];
  bool ok;
  {
    cd V[16], Q[4];
    ok = modes_layer<GENERAL>(L[0], w, kx, ky, V, Q);
#pragma unroll
    for (int i = 0; i < 16; ++i) Vp.set(i, V[i]);
    if (A.V1)
#pragma unroll
      for (int i = 0; i < 16; ++i) A.V1[i * A.ntot + g] = V[i];
    if (A.Q1)
#pragma unroll
      for (int i = 0; i < 4; ++i) A.Q1[i * A.ntot + g] = Q[i];
    prop4(Q, L[0].d, p);
  }
  uint64_t bits = 0;
  for (int j = 1; j < nl; ++j) {
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
      if (j == nl - 1) {
        if (A.VN)
#pragma unroll
          for (int i = 0; i < 16; ++i) A.VN[i * A.ntot + g] = V[i];
        if (A.QN)
#pragma unroll
          for (int i = 0; i < 4; ++i) A.QN[i * A.ntot + g] = Q[i];
      }
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
    cd St[16];
    if (j == 1) redheffer_ldiag(p, Si, St);
    else redheffer(acc, Si, St);
    if (j == nl - 1) {
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = St[i];
    } else {
      prop4(Q, L[j].d, p);
      redheffer_rdiag(St, p, acc);
    }
  }
  bool fin = true;
#pragma unroll
  for (int i = 0; i < 16; ++i) fin = fin && isfin(acc[i]);
  // a fast solve that met a singular interface is invalid until its redo
  const bool valid = ok && fin && !redo;
#pragma unroll
  for (int i = 0; i < 16; ++i) A.S[i * A.ntot + g] = valid ? acc[i] : mk(0.0);
  if (A.flags) A.flags[g] = (unsigned char)((valid ? 1u : 0u) | (redo ? 4u : 0u) | (ok ? 0u : 8u));
}

template <bool GENERAL, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_batch_persistent(BatchArgs A) {
  __shared__ cd vprev[WPB * 16 * 64];
  const int lane = threadIdx.x & 63;
  const LaneMat Vp{vprev + (threadIdx.x >> 6) * 16 * 64, lane};
  for (;;) {
    unsigned int it = 0;
    if (lane == 0) it = atomicAdd(A.counter, 1u);
    it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
    if (it >= (unsigned int)A.n_items) break;
    const int s = __builtin_amdgcn_readfirstlane(A.item_stack[it]);
    const int first = __builtin_amdgcn_readfirstlane(A.item_first[it]);
    const int nl = __builtin_amdgcn_readfirstlane(A.stack_nl[s]);
    const LayerDev *L = A.L + A.stack_l0[s];
    const int64_t n0 = A.stack_n0[s], nn = A.stack_nn[s];
    for (int k = first + lane; k < first + A.nodes_per_item && k < nn; k += 64) {
      bool redo = false;
      stack_node<GENERAL, false>(A, L, nl, n0 + k, Vp, redo);
      if (__builtin_expect(redo, 0)) redo_mark(A.redo, n0 + k);
    }
  }
}

// The nodes whose interface solve was exactly singular, solved again with the
// reference's pivoted-QR fallback (mod_ScatteringMatrix.jl:35-44); the marks
// are cleared for the next launch or graph replay.
template <bool GENERAL>
__global__ __launch_bounds__(64) void k_batch_redo(BatchArgs A, int n_stacks) {
  if (!redo_any(A.redo)) return;
  __shared__ cd vprev[16 * 64];
  const LaneMat Vp{vprev, (int)threadIdx.x};
  for (int64_t g = blockIdx.x * 64 + threadIdx.x; g < A.ntot; g += (int64_t)gridDim.x * 64) {
    if (!redo_take(A.redo, g)) continue;
    int lo = 0, hi = n_stacks - 1;  // the stack of node g: the last with stack_n0 <= g
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (A.stack_n0[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
    bool again = false;
    stack_node<GENERAL, true>(A, A.L + A.stack_l0[lo], A.stack_nl[lo], g, Vp, again);
  }
  redo_finish(A.redo);
}

__global__ void k_batch_reset(unsigned int *counter) {
  if (threadIdx.x == 0) *counter = 0u;
}

template <bool GENERAL>
static hipError_t launch_g(const BatchArgs &A, int wpb, int blocks, hipStream_t s) {
  switch (wpb) {
    case 1: hipLaunchKernelGGL((k_batch_persistent<GENERAL, 1>), dim3(blocks), dim3(64), 0, s, A); break;
    case 2: hipLaunchKernelGGL((k_batch_persistent<GENERAL, 2>), dim3(blocks), dim3(128), 0, s, A); break;
    case 4: hipLaunchKernelGGL((k_batch_persistent<GENERAL, 4>), dim3(blocks), dim3(256), 0, s, A); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch(const BatchLaunch &B, hipStream_t s) {
  BatchArgs A;
  A.L = B.L;
  A.stack_l0 = B.stack_l0;
  A.stack_nl = B.stack_nl;
  A.stack_n0 = B.stack_n0;
  A.stack_nn = B.stack_nn;
  A.item_stack = B.item_stack;
  A.item_first = B.item_first;
  A.n_items = B.n_items;
  A.nodes_per_item = B.nodes_per_item;
  A.omega = B.omega;
  A.kx = B.kx;
  A.ky = B.ky;
  A.S = B.S;
  A.V1 = B.V1;
  A.Q1 = B.Q1;
  A.VN = B.VN;
  A.QN = B.QN;
  A.flags = B.flags;
  A.ntot = B.ntot;
  A.counter = B.counter;
  A.redo = B.redo;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, B.counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = B.general ? launch_g<true>(A, B.waves_per_block, B.blocks, s) : launch_g<false>(A, B.waves_per_block, B.blocks, s);
  if (e != hipSuccess) return e;
  if (B.general) hipLaunchKernelGGL(k_batch_redo<true>, dim3(kRedoBlocks), dim3(64), 0, s, A, B.n_stacks);
  else hipLaunchKernelGGL(k_batch_redo<false>, dim3(kRedoBlocks), dim3(64), 0, s, A, B.n_stacks);
  return hipGetLastError();
}

}  // namespace ytf
