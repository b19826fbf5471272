// Host executor behind the C-ABI: one workspace = one device, one stream,
// device + pinned storage sized at Reserve, level-by-level kernel launches.
// Mirrors the behaviour of the reference executor
// (src/gpu_backend_impl.inc:1636-2403 and 2405-3130), redesigned around
// renamed path slots (no tape copies) and runtime precision.
#ifndef TREE_HMM_AMD_WORKSPACE_H_
#define TREE_HMM_AMD_WORKSPACE_H_

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include "launch.h"
#include "plan.h"
#include "tree_hmm_amd.h"

namespace thmm {

// Extra reserve flag (no reference counterpart): skip pinned host staging
// and workspace-owned output buffers; only the device-resident entry points
// may then be used. Keeps multi-GB edge-marginal outputs single-copy.
constexpr unsigned kReserveDeviceResident = 8u;
// Extra reserve flag: device-resident entry points replay captured hipGraphs.
constexpr unsigned kReserveGraph = 16u;
// Extra reserve flag: the device-resident marginal entry points write edge
// marginals (and edge gradients) as IEEE binary16. Needs
// kReserveDeviceResident and THMM_RESERVE_MARGINALS.
constexpr unsigned kReserveHalfEdgeMarginals = 32u;
static_assert(kReserveDeviceResident == THMM_RESERVE_DEVICE_RESIDENT && kReserveGraph == THMM_RESERVE_GRAPH &&
                  kReserveHalfEdgeMarginals == THMM_RESERVE_HALF_EDGE_MARGINALS,
              "reserve flags of the C interface");

class Workspace {
 public:
  Workspace();
  ~Workspace();
  Workspace(const Workspace &) = delete;
  Workspace &operator=(const Workspace &) = delete;

  void Reserve(const Plan &plan, const void *identity, std::size_t states, std::size_t batch,
               Real real, unsigned flags, int device);
  void ReserveCategorical(const Plan &plan, const void *identity, std::size_t states,
                          std::size_t batch, std::size_t categories, const std::uint32_t *observation_nodes,
                          std::size_t observation_count, Real real, unsigned flags, int device);

  void Inputs(std::size_t batch, void **nodes, std::size_t *node_count, void **edges,
              std::size_t *edge_count);
  void CategoricalInputs(std::size_t batch, std::uint8_t **observations, std::size_t *observation_count,
                         void **root, void **emissions, std::size_t *emission_count, void **edges,
                         std::size_t *edge_count);
  void Uniforms(std::size_t batch, void **uniforms, std::size_t *count);

  void Partition(const thmm_dense_model &model, bool log_domain, const void **values, std::size_t *count,
                 thmm_timings *timings);
  void PartitionCategorical(const thmm_categorical_model &model, thmm_categorical_update update,
                            bool log_domain, const void **values, std::size_t *count, thmm_timings *timings);
  void Maximum(const thmm_dense_model &model, const void **log_weights, const std::uint32_t **states,
               thmm_timings *timings);
  void MaximumCategorical(const thmm_categorical_model &model, thmm_categorical_update update,
                          const void **log_weights, const std::uint32_t **states, thmm_timings *timings);
  void Sample(const thmm_dense_model &model, const void *uniforms, std::size_t uniform_count,
              const std::uint32_t **states, thmm_timings *timings);
  void SampleCategorical(const thmm_categorical_model &model, const void *uniforms, std::size_t uniform_count,
                         thmm_categorical_update update, const std::uint32_t **states, thmm_timings *timings);
  void Marginals(const thmm_dense_model &model, const void **log_partitions, const void **nodes,
                 const void **edges, thmm_timings *timings);
  void MarginalsCategorical(const thmm_categorical_model &model, thmm_categorical_update update,
                            const void **log_partitions, const void **nodes, const void **edges,
                            thmm_timings *timings);
  void LogPartitionsF64(const double **values, std::size_t *count);
  void CategoricalLogZForward(const thmm_categorical_model &model, thmm_categorical_update update,
                              const double **log_z, std::size_t *count);
  // The forward half with the log partitions summed over the batch on the
  // device: sums [roots], FP64.
  void CategoricalLogLikelihood(const thmm_categorical_model &model, thmm_categorical_update update,
                                const double **sums, std::size_t *count);
  // Writes the categorical workspace's device edge factors as expm(Q t_e)
  // (reversible Q with stationary law pi); needs resident observations and
  // factors, drops a held forward pass.
  void SetBranchLengths(const double *rate_matrix, const double *stationary, const double *branch_lengths,
                        std::size_t edge_count);
  void BranchLengthBackward(const double *grad_log_z, const double *rate_matrix, const double **grad,
                            std::size_t *count);

  void LogPartitionDevice(const void *nodes, const void *edges, std::size_t batch, double *log_z, void *stream);
  // With `timings` non-null, every launch is bracketed by an event pair and
  // the call blocks to report the durations.
  void MarginalsDevice(const void *nodes, const void *edges, std::size_t batch, double *log_z, void *node_out,
                       void *edge_out, void *stream, thmm_launch_timing *timings = nullptr,
                       std::size_t capacity = 0, std::size_t *count = nullptr);
  void LogZForwardDevice(const void *log_nodes, const void *log_edges, std::size_t batch, double *log_z,
                         void *stream);
  void LogZBackwardDevice(const double *grad, void *grad_nodes, void *grad_edges, bool reduce_edges,
                          void *stream);
  // MAP and posterior sampling on device-resident dense potentials; the
  // outputs are device buffers, written without a host synchronisation.
  void MaximumDevice(const void *nodes, const void *edges, std::size_t batch, double *log_weights,
                     std::uint32_t *states, void *stream);
  void SampleDevice(const void *nodes, const void *edges, const void *uniforms, std::size_t batch, double *log_z,
                    std::uint32_t *states, std::uint32_t *failure, void *stream);
  void Synchronize();

  // Path-storage budget for the next reservations (0: none) and, after a
  // reservation, the tree groups it runs in (0: not grouped) and the bytes
  // of path storage it holds.
  void SetPathBudget(std::size_t bytes);
  void PathStorage(std::size_t *groups, std::size_t *bytes) const;

  // Launch variants (see Tuning in launch.h).
  void SetTuning(const Tuning &t);
  Tuning GetTuning() const;
  // Times the variants on the caller's inputs and keeps the fastest; returns
  // the best measured milliseconds per marginal pass.
  double AutotuneMarginalsDevice(const void *nodes, const void *edges, std::size_t batch, double *log_z,
                                 void *node_out, void *edge_out, void *stream);

  struct Impl;

 private:
  std::unique_ptr<Impl> impl_;
};

}  // namespace thmm

#endif  // TREE_HMM_AMD_WORKSPACE_H_
