// node_general.hip -- launches of the Green k-node kernel for every other
// (source layer, observation layer) pair: same layer inside the stack
// (CASE_SAME), below (CASE_RIGHT) or above (CASE_LEFT) the source.
//
// Built at -O3 like every other file.  An earlier formulation of these cases,
// which kept full 4x4 S-matrix accumulators live through the layer loop and
// spilled heavily, gave wrong results on gfx950 (a host build of the same code
// was exact).  The current two-pass cascade (cascade_general in node_core.hpp)
// keeps its carried blocks in LDS; the host build of this source under
// AddressSanitizer and UndefinedBehaviorSanitizer, with every LDS slot
// poisoned before each node, finds no out-of-range slot, no slot read before
// it is written and no undefined behaviour in any case
// (tests/test_host_node_sanitized.py).
#include "node_core.hpp"

namespace ytf {

template <int CASE, bool GENERAL>
static void launch(const NodeParams &P, int blocks, int groups, hipStream_t stream) {
  hipLaunchKernelGGL((node_kernel<CASE, GENERAL>), dim3(blocks, groups), dim3(64), 0, stream, P);
  if (P.redo) hipLaunchKernelGGL((node_kernel_redo<CASE, GENERAL>), dim3(kRedoBlocks), dim3(64), 0, stream, P);
}

template <bool GENERAL>
static void launch_case(const NodeParams &P, int blocks, int groups, hipStream_t stream) {
  if (P.obs == P.src) launch<CASE_SAME, GENERAL>(P, blocks, groups, stream);
  else if (P.obs > P.src) launch<CASE_RIGHT, GENERAL>(P, blocks, groups, stream);
  else launch<CASE_LEFT, GENERAL>(P, blocks, groups, stream);
}

hipError_t launch_node_general(const NodeParams &P, int blocks, hipStream_t stream, int pair_groups) {
  // the general 4x4 eigen-solver is compiled only into the variant for stacks
  // that need it (a layer with xz/yz coupling in eps or mu); its registers
  // would otherwise be reserved by every launch
  if (P.needs_general) launch_case<true>(P, blocks, pair_groups, stream);
  else launch_case<false>(P, blocks, pair_groups, stream);
  return hipGetLastError();
}

}  // namespace ytf
