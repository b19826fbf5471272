// Workspace-free log partition (thmm_log_partition_device_static): the
// contraction state lives in a file-scope device array of each GPU.
#ifndef TREE_HMM_AMD_STATIC_SCRATCH_H_
#define TREE_HMM_AMD_STATIC_SCRATCH_H_

#include <cstddef>

#include "plan.h"

namespace thmm {

// Bytes of contraction state the call needs (ops, node and path state in
// 256-byte aligned pieces) for this plan, state count, batch and precision.
std::size_t StaticScratchBytes(const Plan &plan, std::size_t states, std::size_t batch, Real real);

// Scaled forward pass and log partitions, as thmm_log_partition_device runs
// them through the plan's levels, with every buffer beyond the caller's
// inputs and outputs in the device array. Allocates no device memory; calls
// that overlap on one GPU share (and race on) the array.
void LogPartitionStatic(const Plan &plan, const RenamedOps &ops, std::size_t states, std::size_t batch, Real real,
                        const void *nodes, const void *edges, double *log_z, void *stream, int device);

}  // namespace thmm

#endif  // TREE_HMM_AMD_STATIC_SCRATCH_H_
