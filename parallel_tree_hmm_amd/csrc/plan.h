// Rake-compress contraction planner (host C++).
//
// The reference takes its plans from the un-vendored btrc library
// (MODULE.bazel:10,15-19); only its call sites are visible
// (src/inference.cc:155-214,888-937; src/gpu_backend_impl.inc:39-42,2245).
// This planner produces the same kinds of ops with the same field meaning,
// grouped into write-disjoint levels:
//
//   per round:  [rakes] [combination stage 1..s] [absorptions] [compressions]
//
// * Rake:        every live non-root leaf is raked into its live parent over
//                the edge slot that currently joins them; one new branch each.
// * Combination: a parent's branches of this round are reduced pairwise,
//                one stage per halving (branch_reduction_stages).
// * Absorption:  each parent absorbs its one remaining branch.
// * Compression: maximal chains of unary live nodes are found and every other
//                node (starting at the chain head) is spliced out; the
//                composed path is written into the middle's parent-edge slot
//                (left_edge), which then joins parent and child.
//
// Each round removes every leaf and at least half of every unary chain, so
// the round count is O(log N) for any topology.
#ifndef TREE_HMM_AMD_PLAN_H_
#define TREE_HMM_AMD_PLAN_H_

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "launch.h"
#include "tree_hmm_amd.h"

namespace thmm {

using Index = std::uint32_t;

struct Plan {
  std::size_t num_nodes = 0;
  std::size_t num_edges = 0;
  std::size_t num_branches = 0;
  std::vector<Index> roots;
  std::vector<Index> edge_parents;
  std::vector<Index> edge_children;
  std::vector<thmm_rake> rakes;
  std::vector<thmm_branch_combination> combinations;
  std::vector<thmm_branch_absorption> absorptions;
  std::vector<thmm_compression> compressions;
  std::vector<thmm_primitive_batch> batches;
  std::vector<Index> round_batch_offsets;  // [rounds + 1]

  std::size_t num_compressions() const { return compressions.size(); }
  std::size_t num_rounds() const {
    return round_batch_offsets.empty() ? 0 : round_batch_offsets.size() - 1;
  }
  thmm_plan_arrays Arrays() const;
};

// The plan's ops with renamed path slots (launch.h): slots 0..E-1 are the
// caller's edge potentials and compression c writes the fresh slot E + c
// instead of overwriting its left edge, as src/inference.cc:191-214 does.
struct RenamedOps {
  std::vector<DeviceRake> rakes;
  std::vector<DeviceCombine> combines;
  std::vector<DeviceAbsorb> absorbs;
  std::vector<DeviceCompress> compresses;
};
RenamedOps RenameSlots(const Plan &plan);

// Builds a plan from a parent array (-1 = root). Throws invalid_argument for
// out-of-range parents or cycles and length_error beyond 2^32 - 2 nodes.
Plan MakePlan(const std::int64_t *parents, std::size_t num_nodes);

// The same plan as MakePlan, built on device `device` (device_plan.hip).
// `parents` is a host array, or with parents_on_device a device array on
// that device. Same errors as MakePlan; at most 2^31 - 1 nodes.
Plan MakePlanOnDevice(const std::int64_t *parents, std::size_t num_nodes, bool parents_on_device, int device);

// Copies an external plan and verifies it (see Validate).
Plan PlanFromArrays(const thmm_plan_arrays &arrays);

// Checks the invariants the executor relies on; throws invalid_argument with
// a description of the first violation:
//  * edge_parents/edge_children describe a forest whose roots are `roots`;
//  * every non-root node is eliminated exactly once (as a rake leaf or a
//    compression middle), so rakes + compressions == num_edges;
//  * every branch is produced by one rake and consumed once (by a
//    combination as source or by an absorption);
//  * ops inside one batch write disjoint storage and never read what another
//    op of the same batch writes;
//  * replaying the batches in order keeps every op's operands live.
void Validate(const Plan &plan);

}  // namespace thmm

#endif  // TREE_HMM_AMD_PLAN_H_
