// node_top.hip -- launch of the Green k-node kernel when the source and the
// observation plane are both in the first layer (CASE_TOP, the source above
// the stack), compiled at -O3.
#include "node_core.hpp"

namespace ytf {

hipError_t launch_node_top(const NodeParams &P, int blocks, hipStream_t stream, int pair_groups) {
  const dim3 grid(blocks, pair_groups);
  if (P.needs_general) hipLaunchKernelGGL((node_kernel<CASE_TOP, true>), grid, dim3(64), 0, stream, P);
  else hipLaunchKernelGGL((node_kernel<CASE_TOP, false>), grid, dim3(64), 0, stream, P);
  if (P.redo) {
    if (P.needs_general) hipLaunchKernelGGL((node_kernel_redo<CASE_TOP, true>), dim3(kRedoBlocks), dim3(64), 0, stream, P);
    else hipLaunchKernelGGL((node_kernel_redo<CASE_TOP, false>), dim3(kRedoBlocks), dim3(64), 0, stream, P);
  }
  return hipGetLastError();
}

}  // namespace ytf
