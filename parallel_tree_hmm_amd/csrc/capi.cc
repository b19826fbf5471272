// extern "C" entry points of libtree_hmm_amd.so (declared in
// include/tree_hmm_amd.h). Each one catches the C++ exception the executor
// throws and returns the matching status; the message is kept per thread.
#include <hip/hip_runtime_api.h>

#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

#include "expm.h"
#include "plan.h"
#include "static_scratch.h"
#include "tree_hmm_amd.h"
#include "workspace.h"

struct thmm_plan {
  thmm::Plan plan;
  // Renamed ops for the workspace-free entry point (formed once here, so
  // that entry point allocates nothing).
  thmm::RenamedOps renamed;
};

struct thmm_workspace {
  thmm::Workspace workspace;
};

namespace {

thread_local std::string g_last_error;

template <class F> thmm_status Guard(F body) {
  try {
    body();
    g_last_error.clear();
    return THMM_OK;
  } catch (const std::invalid_argument &e) {
    g_last_error = e.what();
    return THMM_INVALID_ARGUMENT;
  } catch (const std::domain_error &e) {
    g_last_error = e.what();
    return THMM_DOMAIN_ERROR;
  } catch (const std::length_error &e) {
    g_last_error = e.what();
    return THMM_LENGTH_ERROR;
  } catch (const std::logic_error &e) {
    g_last_error = e.what();
    return THMM_LOGIC_ERROR;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return THMM_RUNTIME_ERROR;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return THMM_RUNTIME_ERROR;
  }
}

void Require(bool condition, const char *message) {
  if (!condition) throw std::invalid_argument(message);
}

thmm::Real RealOf(thmm_precision precision) {
  if (precision == THMM_F32) return thmm::Real::kF32;
  if (precision == THMM_F64) return thmm::Real::kF64;
  throw std::invalid_argument("unknown precision");
}

thmm::Workspace &W(thmm_workspace *workspace) {
  Require(workspace != nullptr, "workspace is null");
  return workspace->workspace;
}

thmm_tuning ToC(const thmm::Tuning &t, double ms) {
  return {t.compress_items_forward, t.compress_items_reverse, t.rake_block, t.gather_chunk, t.compress_depth,
          t.rake_log_edges, t.persistent, t.input_leaves, ms};
}

}  // namespace

extern "C" {

const char *thmm_last_error(void) { return g_last_error.c_str(); }

thmm_status thmm_plan_create(const int64_t *parents, size_t num_nodes, thmm_plan **out) {
  return Guard([&] {
    Require(out != nullptr, "output pointer is null");
    Require(parents != nullptr || num_nodes == 0, "parents are null");
    auto *plan = new thmm_plan{thmm::MakePlan(parents, num_nodes), {}};
    plan->renamed = thmm::RenameSlots(plan->plan);
    *out = plan;
  });
}

thmm_status thmm_plan_create_device(const int64_t *parents, size_t num_nodes, int parents_on_device, int device,
                                    thmm_plan **out) {
  return Guard([&] {
    Require(out != nullptr, "output pointer is null");
    Require(parents != nullptr || num_nodes == 0, "parents are null");
    auto *plan = new thmm_plan{thmm::MakePlanOnDevice(parents, num_nodes, parents_on_device != 0, device), {}};
    plan->renamed = thmm::RenameSlots(plan->plan);
    *out = plan;
  });
}

thmm_status thmm_plan_create_from_arrays(const thmm_plan_arrays *arrays, thmm_plan **out) {
  return Guard([&] {
    Require(out != nullptr && arrays != nullptr, "argument is null");
    auto *plan = new thmm_plan{thmm::PlanFromArrays(*arrays), {}};
    plan->renamed = thmm::RenameSlots(plan->plan);
    *out = plan;
  });
}

thmm_status thmm_plan_get_arrays(const thmm_plan *plan, thmm_plan_arrays *out) {
  return Guard([&] {
    Require(plan != nullptr && out != nullptr, "argument is null");
    *out = plan->plan.Arrays();
  });
}

void thmm_plan_destroy(thmm_plan *plan) { delete plan; }

int thmm_available(void) {
  int count = 0;
  return hipGetDeviceCount(&count) == hipSuccess && count > 0 ? 1 : 0;
}

thmm_status thmm_device_description(int device, char *buffer, size_t size) {
  return Guard([&] {
    Require(buffer != nullptr && size > 0, "buffer is empty");
    hipDeviceProp_t properties{};
    const hipError_t status = hipGetDeviceProperties(&properties, device);
    if (status != hipSuccess)
      throw std::runtime_error(std::string("GPU runtime GetDeviceProperties: ") + hipGetErrorString(status));
    const std::string text = std::string(properties.name) + " (" + properties.gcnArchName + ")";
    std::strncpy(buffer, text.c_str(), size - 1);
    buffer[size - 1] = '\0';
  });
}

thmm_status thmm_workspace_create(thmm_workspace **out) {
  return Guard([&] {
    Require(out != nullptr, "output pointer is null");
    *out = new thmm_workspace{};
  });
}

void thmm_workspace_destroy(thmm_workspace *workspace) { delete workspace; }

thmm_status thmm_reserve(thmm_workspace *workspace, const thmm_plan *plan, const void *identity, size_t states,
                         size_t batch_capacity, thmm_precision precision, unsigned flags, int device) {
  return Guard([&] {
    Require(plan != nullptr, "plan is null");
    W(workspace).Reserve(plan->plan, identity, states, batch_capacity, RealOf(precision), flags, device);
  });
}

thmm_status thmm_reserve_categorical(thmm_workspace *workspace, const thmm_plan *plan, const void *identity,
                                     size_t states, size_t batch_capacity, size_t categories,
                                     const uint32_t *observation_nodes, size_t num_observation_nodes,
                                     thmm_precision precision, unsigned flags, int device) {
  return Guard([&] {
    Require(plan != nullptr, "plan is null");
    W(workspace).ReserveCategorical(plan->plan, identity, states, batch_capacity, categories, observation_nodes,
                                    num_observation_nodes, RealOf(precision), flags, device);
  });
}

thmm_status thmm_inputs(thmm_workspace *workspace, size_t batch, void **node_potentials, size_t *node_count,
                        void **edge_potentials, size_t *edge_count) {
  return Guard([&] {
    Require(node_potentials && node_count && edge_potentials && edge_count, "output pointer is null");
    W(workspace).Inputs(batch, node_potentials, node_count, edge_potentials, edge_count);
  });
}

thmm_status thmm_categorical_inputs(thmm_workspace *workspace, size_t batch, uint8_t **observations,
                                    size_t *observation_count, void **root_potential, void **emission_potentials,
                                    size_t *emission_count, void **edge_potentials, size_t *edge_count) {
  return Guard([&] {
    Require(observations && observation_count && root_potential && emission_potentials && emission_count &&
                edge_potentials && edge_count,
            "output pointer is null");
    W(workspace).CategoricalInputs(batch, observations, observation_count, root_potential, emission_potentials,
                                   emission_count, edge_potentials, edge_count);
  });
}

thmm_status thmm_uniforms(thmm_workspace *workspace, size_t batch, void **uniforms, size_t *count) {
  return Guard([&] {
    Require(uniforms && count, "output pointer is null");
    W(workspace).Uniforms(batch, uniforms, count);
  });
}

thmm_status thmm_partition(thmm_workspace *workspace, const thmm_dense_model *model, int log_domain,
                           const void **values, size_t *count, thmm_timings *timings) {
  return Guard([&] {
    Require(model && values && count, "argument is null");
    W(workspace).Partition(*model, log_domain != 0, values, count, timings);
  });
}

thmm_status thmm_partition_categorical(thmm_workspace *workspace, const thmm_categorical_model *model,
                                       thmm_categorical_update update, int log_domain, const void **values,
                                       size_t *count, thmm_timings *timings) {
  return Guard([&] {
    Require(model && values && count, "argument is null");
    W(workspace).PartitionCategorical(*model, update, log_domain != 0, values, count, timings);
  });
}

thmm_status thmm_maximum(thmm_workspace *workspace, const thmm_dense_model *model, const void **log_weights,
                         const uint32_t **states, thmm_timings *timings) {
  return Guard([&] {
    Require(model && log_weights && states, "argument is null");
    W(workspace).Maximum(*model, log_weights, states, timings);
  });
}

thmm_status thmm_maximum_categorical(thmm_workspace *workspace, const thmm_categorical_model *model,
                                     thmm_categorical_update update, const void **log_weights,
                                     const uint32_t **states, thmm_timings *timings) {
  return Guard([&] {
    Require(model && log_weights && states, "argument is null");
    W(workspace).MaximumCategorical(*model, update, log_weights, states, timings);
  });
}

thmm_status thmm_sample(thmm_workspace *workspace, const thmm_dense_model *model, const void *uniforms,
                        size_t uniform_count, const uint32_t **states, thmm_timings *timings) {
  return Guard([&] {
    Require(model && states, "argument is null");
    W(workspace).Sample(*model, uniforms, uniform_count, states, timings);
  });
}

thmm_status thmm_sample_categorical(thmm_workspace *workspace, const thmm_categorical_model *model,
                                    const void *uniforms, size_t uniform_count, thmm_categorical_update update,
                                    const uint32_t **states, thmm_timings *timings) {
  return Guard([&] {
    Require(model && states, "argument is null");
    W(workspace).SampleCategorical(*model, uniforms, uniform_count, update, states, timings);
  });
}

thmm_status thmm_marginals(thmm_workspace *workspace, const thmm_dense_model *model, const void **log_partitions,
                           const void **nodes, const void **edges, thmm_timings *timings) {
  return Guard([&] {
    Require(model && log_partitions && nodes && edges, "argument is null");
    W(workspace).Marginals(*model, log_partitions, nodes, edges, timings);
  });
}

thmm_status thmm_marginals_categorical(thmm_workspace *workspace, const thmm_categorical_model *model,
                                       thmm_categorical_update update, const void **log_partitions,
                                       const void **nodes, const void **edges, thmm_timings *timings) {
  return Guard([&] {
    Require(model && log_partitions && nodes && edges, "argument is null");
    W(workspace).MarginalsCategorical(*model, update, log_partitions, nodes, edges, timings);
  });
}

thmm_status thmm_result_log_partitions_f64(thmm_workspace *workspace, const double **values, size_t *count) {
  return Guard([&] {
    Require(values && count, "argument is null");
    W(workspace).LogPartitionsF64(values, count);
  });
}

thmm_status thmm_categorical_logz_forward(thmm_workspace *workspace, const thmm_categorical_model *model,
                                          thmm_categorical_update update, const double **log_z, size_t *count) {
  return Guard([&] {
    Require(model && log_z && count, "argument is null");
    W(workspace).CategoricalLogZForward(*model, update, log_z, count);
  });
}

thmm_status thmm_categorical_log_likelihood(thmm_workspace *workspace, const thmm_categorical_model *model,
                                            thmm_categorical_update update, const double **sums, size_t *count) {
  return Guard([&] {
    Require(model && sums && count, "argument is null");
    W(workspace).CategoricalLogLikelihood(*model, update, sums, count);
  });
}

thmm_status thmm_branch_length_backward(thmm_workspace *workspace, const double *grad_log_z,
                                        const double *rate_matrix, const double **grad, size_t *count) {
  return Guard([&] {
    Require(grad && count, "argument is null");
    W(workspace).BranchLengthBackward(grad_log_z, rate_matrix, grad, count);
  });
}

thmm_status thmm_expm_reversible(const double *rate_matrix, const double *stationary, size_t states,
                                 const double *branch_lengths, size_t edges, double *out) {
  return Guard([&] {
    Require(edges == 0 || out != nullptr, "output is null");
    const thmm::ReversibleRates rates = thmm::DecomposeReversible(rate_matrix, stationary, states);
    thmm::CheckBranchLengths(branch_lengths, edges);
    thmm::ExpmReversible(rates, branch_lengths, edges, out);
  });
}

thmm_status thmm_categorical_set_branch_lengths(thmm_workspace *workspace, const double *rate_matrix,
                                                const double *stationary, const double *branch_lengths,
                                                size_t edge_count) {
  return Guard([&] { W(workspace).SetBranchLengths(rate_matrix, stationary, branch_lengths, edge_count); });
}

thmm_status thmm_get_tuning(const thmm_workspace *workspace, thmm_tuning *out) {
  return Guard([&] {
    Require(workspace && out, "argument is null");
    *out = ToC(workspace->workspace.GetTuning(), 0.0);
  });
}

thmm_status thmm_set_tuning(thmm_workspace *workspace, const thmm_tuning *tuning) {
  return Guard([&] {
    Require(tuning, "argument is null");
    thmm::Tuning t;
    t.compress_items_forward = tuning->compress_items_forward;
    t.compress_items_reverse = tuning->compress_items_reverse;
    t.rake_block = tuning->rake_block;
    t.gather_chunk = tuning->gather_chunk;
    t.compress_depth = tuning->compress_depth;
    t.rake_log_edges = tuning->rake_log_edges;
    t.persistent = tuning->persistent;
    t.input_leaves = tuning->input_leaves;
    W(workspace).SetTuning(t);
  });
}

thmm_status thmm_autotune_marginals_device(thmm_workspace *workspace, const void *d_node_potentials,
                                           const void *d_edge_potentials, size_t batch, double *d_log_z,
                                           void *d_node_marginals, void *d_edge_marginals, void *stream,
                                           thmm_tuning *chosen) {
  return Guard([&] {
    Require(chosen, "argument is null");
    thmm::Workspace &w = W(workspace);
    const double ms = w.AutotuneMarginalsDevice(d_node_potentials, d_edge_potentials, batch, d_log_z,
                                                d_node_marginals, d_edge_marginals, stream);
    *chosen = ToC(w.GetTuning(), ms);
  });
}

thmm_status thmm_log_partition_device(thmm_workspace *workspace, const void *d_node_potentials,
                                      const void *d_edge_potentials, size_t batch, double *d_log_z, void *stream) {
  return Guard([&] { W(workspace).LogPartitionDevice(d_node_potentials, d_edge_potentials, batch, d_log_z, stream); });
}

thmm_status thmm_set_path_budget(thmm_workspace *workspace, size_t bytes) {
  return Guard([&] { W(workspace).SetPathBudget(bytes); });
}

thmm_status thmm_path_storage(const thmm_workspace *workspace, size_t *groups, size_t *bytes) {
  return Guard([&] {
    Require(workspace != nullptr && groups != nullptr && bytes != nullptr, "argument is null");
    workspace->workspace.PathStorage(groups, bytes);
  });
}

thmm_status thmm_static_scratch_bytes(const thmm_plan *plan, size_t states, size_t batch,
                                      thmm_precision precision, size_t *bytes) {
  return Guard([&] {
    Require(plan != nullptr && bytes != nullptr, "argument is null");
    *bytes = thmm::StaticScratchBytes(plan->plan, states, batch, RealOf(precision));
  });
}

thmm_status thmm_log_partition_device_static(const thmm_plan *plan, size_t states, size_t batch,
                                             thmm_precision precision, const void *d_node_potentials,
                                             const void *d_edge_potentials, double *d_log_z, void *stream,
                                             int device) {
  return Guard([&] {
    Require(plan != nullptr, "plan is null");
    thmm::LogPartitionStatic(plan->plan, plan->renamed, states, batch, RealOf(precision), d_node_potentials,
                             d_edge_potentials, d_log_z, stream, device);
  });
}

thmm_status thmm_marginals_device(thmm_workspace *workspace, const void *d_node_potentials,
                                  const void *d_edge_potentials, size_t batch, double *d_log_z,
                                  void *d_node_marginals, void *d_edge_marginals, void *stream) {
  return Guard([&] {
    W(workspace).MarginalsDevice(d_node_potentials, d_edge_potentials, batch, d_log_z, d_node_marginals,
                                 d_edge_marginals, stream);
  });
}

thmm_status thmm_logz_forward_device(thmm_workspace *workspace, const void *d_log_node_potentials,
                                     const void *d_log_edge_potentials, size_t batch, double *d_log_z,
                                     void *stream) {
  return Guard([&] {
    W(workspace).LogZForwardDevice(d_log_node_potentials, d_log_edge_potentials, batch, d_log_z, stream);
  });
}

thmm_status thmm_logz_backward_device(thmm_workspace *workspace, const double *d_grad_log_z, void *d_grad_nodes,
                                      void *d_grad_edges, int reduce_edges_over_batch, void *stream) {
  return Guard([&] {
    W(workspace).LogZBackwardDevice(d_grad_log_z, d_grad_nodes, d_grad_edges, reduce_edges_over_batch != 0, stream);
  });
}

thmm_status thmm_maximum_device(thmm_workspace *workspace, const void *d_node_potentials,
                                const void *d_edge_potentials, size_t batch, double *d_log_weights,
                                uint32_t *d_states, void *stream) {
  return Guard([&] {
    W(workspace).MaximumDevice(d_node_potentials, d_edge_potentials, batch, d_log_weights, d_states, stream);
  });
}

thmm_status thmm_sample_device(thmm_workspace *workspace, const void *d_node_potentials, const void *d_edge_potentials,
                               const void *d_uniforms, size_t batch, double *d_log_z, uint32_t *d_states,
                               uint32_t *d_failure, void *stream) {
  return Guard([&] {
    W(workspace).SampleDevice(d_node_potentials, d_edge_potentials, d_uniforms, batch, d_log_z, d_states, d_failure,
                              stream);
  });
}

thmm_status thmm_profile_marginals_device(thmm_workspace *workspace, const void *d_node_potentials,
                                          const void *d_edge_potentials, size_t batch, double *d_log_z,
                                          void *d_node_marginals, void *d_edge_marginals, void *stream,
                                          thmm_launch_timing *timings, size_t capacity, size_t *count) {
  return Guard([&] {
    Require(timings != nullptr && count != nullptr, "timing output is null");
    W(workspace).MarginalsDevice(d_node_potentials, d_edge_potentials, batch, d_log_z, d_node_marginals,
                                 d_edge_marginals, stream, timings, capacity, count);
  });
}

thmm_status thmm_synchronize(thmm_workspace *workspace) {
  return Guard([&] { W(workspace).Synchronize(); });
}

}  // extern "C"
