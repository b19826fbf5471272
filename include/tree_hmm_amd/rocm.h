// Header-only C++20 layer that rebuilds the reference accelerator API
// (include/tree_hmm/accelerator.h, include/tree_hmm/rocm.h and
// include/tree_hmm/detail/gpu_backend_api.inc of joaospinto/parallel_tree_hmm)
// on top of the C-ABI in tree_hmm_amd.h. Status codes become the same
// exception classes the reference throws.
//
// Plans: the reference takes `const btrc::Plan &`. This layer accepts any
// plan type P that exposes btrc's accessors (num_nodes, num_edges,
// num_branches, root/roots, edge_parents, edge_children, rakes,
// branch_combinations, branch_absorptions, compressions, primitive_batches),
// flattening it once per Reserve through thmm_plan_create_from_arrays. Where
// btrc is absent, tree_hmm_amd::Plan (built by tree_hmm_amd::MakePlan from a
// parent array) is the plan type; defining TREE_HMM_AMD_PROVIDE_BTRC also
// names it btrc::Plan / btrc::MakePlan so code written against the reference
// compiles unchanged.
//
// Scalar follows the reference's compile-time switch (include/tree_hmm/
// scalar.h:6-12): float when TREE_HMM_USE_FLOAT is defined, else double.
#ifndef TREE_HMM_AMD_ROCM_H_
#define TREE_HMM_AMD_ROCM_H_

#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "tree_hmm_amd.h"

namespace tree_hmm_amd {

[[noreturn]] inline void Throw(thmm_status status) {
  const std::string message = thmm_last_error();
  switch (status) {
    case THMM_INVALID_ARGUMENT: throw std::invalid_argument(message);
    case THMM_LOGIC_ERROR: throw std::logic_error(message);
    case THMM_LENGTH_ERROR: throw std::length_error(message);
    case THMM_DOMAIN_ERROR: throw std::domain_error(message);
    default: throw std::runtime_error(message);
  }
}

inline void Check(thmm_status status) {
  if (status != THMM_OK) Throw(status);
}

using Index = std::uint32_t;

// Owning wrapper of a thmm_plan with btrc::Plan-style accessors.
class Plan {
 public:
  explicit Plan(const std::vector<std::int64_t> &parents) {
    thmm_plan *raw = nullptr;
    Check(thmm_plan_create(parents.data(), parents.size(), &raw));
    Adopt(raw);
  }
  explicit Plan(const thmm_plan_arrays &arrays) {
    thmm_plan *raw = nullptr;
    Check(thmm_plan_create_from_arrays(&arrays, &raw));
    Adopt(raw);
  }
  Plan(Plan &&) noexcept = default;
  Plan &operator=(Plan &&) noexcept = default;

  std::size_t num_nodes() const { return a_.num_nodes; }
  std::size_t num_edges() const { return a_.num_edges; }
  std::size_t num_branches() const { return a_.num_branches; }
  std::size_t num_compressions() const { return a_.num_compressions; }
  Index root() const {
    if (a_.num_roots != 1) throw std::logic_error("plan has more than one root");
    return a_.roots[0];
  }
  std::span<const Index> roots() const { return {a_.roots, a_.num_roots}; }
  std::span<const Index> edge_parents() const { return {a_.edge_parents, a_.num_edges}; }
  std::span<const Index> edge_children() const { return {a_.edge_children, a_.num_edges}; }
  std::span<const thmm_rake> rakes() const { return {a_.rakes, a_.num_rakes}; }
  std::span<const thmm_branch_combination> branch_combinations() const {
    return {a_.combinations, a_.num_combinations};
  }
  std::span<const thmm_branch_absorption> branch_absorptions() const { return {a_.absorptions, a_.num_absorptions}; }
  std::span<const thmm_compression> compressions() const { return {a_.compressions, a_.num_compressions}; }
  std::span<const thmm_primitive_batch> primitive_batches() const { return {a_.batches, a_.num_batches}; }

  // Sibling-reduction stages of each round (tests/accelerator_test.h:241).
  struct Round {
    std::vector<thmm_primitive_batch> branch_reduction_stages;
  };
  std::vector<Round> rounds() const {
    std::vector<Round> out(a_.num_rounds);
    for (std::size_t r = 0; r < a_.num_rounds; ++r)
      for (std::uint32_t b = a_.round_batch_offsets[r]; b < a_.round_batch_offsets[r + 1]; ++b)
        if (a_.batches[b].primitive == THMM_BRANCH_COMBINATION) out[r].branch_reduction_stages.push_back(a_.batches[b]);
    return out;
  }

  const thmm_plan *handle() const { return handle_.get(); }
  const thmm_plan_arrays &arrays() const { return a_; }

 private:
  struct Deleter {
    void operator()(thmm_plan *p) const { thmm_plan_destroy(p); }
  };
  void Adopt(thmm_plan *raw) {
    handle_.reset(raw);
    Check(thmm_plan_get_arrays(raw, &a_));
  }
  std::unique_ptr<thmm_plan, Deleter> handle_;
  thmm_plan_arrays a_{};
};

inline Plan MakePlan(const std::vector<std::int64_t> &parents) { return Plan(parents); }

// Flattens a btrc-style plan into a library plan. For tree_hmm_amd::Plan the
// existing handle is reused.
template <class P> struct PlanBinding {
  std::shared_ptr<const thmm_plan> handle;
  static PlanBinding Bind(const P &plan) {
    if constexpr (std::is_same_v<P, Plan>) {
      return {std::shared_ptr<const thmm_plan>(plan.handle(), [](const thmm_plan *) {})};
    } else {
      std::vector<Index> roots{static_cast<Index>(plan.root())};
      std::vector<thmm_rake> rakes;
      for (const auto &r : plan.rakes()) rakes.push_back({r.leaf, r.parent, r.edge, r.branch});
      std::vector<thmm_branch_combination> combinations;
      for (const auto &c : plan.branch_combinations()) combinations.push_back({c.destination, c.source, 0});
      std::vector<thmm_branch_absorption> absorptions;
      for (const auto &a : plan.branch_absorptions()) absorptions.push_back({a.parent, a.branch, 0});
      std::vector<thmm_compression> compressions;
      for (const auto &c : plan.compressions())
        compressions.push_back({c.left_edge, c.middle, c.right_edge, c.parent, c.child, c.tape});
      std::vector<thmm_primitive_batch> batches;
      for (const auto &b : plan.primitive_batches())
        batches.push_back({static_cast<std::uint32_t>(b.primitive), static_cast<std::uint32_t>(b.offset),
                           static_cast<std::uint32_t>(b.count)});
      std::vector<Index> parents(plan.edge_parents().begin(), plan.edge_parents().end());
      std::vector<Index> children(plan.edge_children().begin(), plan.edge_children().end());
      thmm_plan_arrays a{};
      a.num_nodes = plan.num_nodes();
      a.num_edges = plan.num_edges();
      a.num_branches = plan.num_branches();
      a.num_compressions = compressions.size();
      a.num_roots = 1;
      a.num_rakes = rakes.size();
      a.num_combinations = combinations.size();
      a.num_absorptions = absorptions.size();
      a.num_batches = batches.size();
      a.roots = roots.data();
      a.edge_parents = parents.data();
      a.edge_children = children.data();
      a.rakes = rakes.data();
      a.combinations = combinations.data();
      a.absorptions = absorptions.data();
      a.compressions = compressions.data();
      a.batches = batches.data();
      thmm_plan *raw = nullptr;
      Check(thmm_plan_create_from_arrays(&a, &raw));
      return {std::shared_ptr<const thmm_plan>(raw, [](const thmm_plan *p) { thmm_plan_destroy(const_cast<thmm_plan *>(p)); })};
    }
  }
};

}  // namespace tree_hmm_amd

#ifdef TREE_HMM_AMD_PROVIDE_BTRC
namespace btrc {
using Plan = tree_hmm_amd::Plan;
using Index = tree_hmm_amd::Index;
inline Plan MakePlan(const std::vector<std::int64_t> &parents) { return Plan(parents); }
}  // namespace btrc
#define TREE_HMM_AMD_PLAN_TYPE btrc::Plan
#endif

#ifndef TREE_HMM_AMD_PLAN_TYPE
#define TREE_HMM_AMD_PLAN_TYPE tree_hmm_amd::Plan
#endif

namespace tree_hmm {

#ifdef TREE_HMM_USE_FLOAT
using Scalar = float;
inline constexpr const char *kPrecisionName = "FP32";
inline constexpr thmm_precision kAmdPrecision = THMM_F32;
#else
using Scalar = double;
inline constexpr const char *kPrecisionName = "FP64";
inline constexpr thmm_precision kAmdPrecision = THMM_F64;
#endif

using PlanType = TREE_HMM_AMD_PLAN_TYPE;

// include/tree_hmm/accelerator.h:13-146
struct BatchedModelView {
  const PlanType &plan;
  std::size_t states;
  std::size_t batch;
  std::span<const Scalar> node_potentials;
  std::span<const Scalar> edge_potentials;
};

struct MutableBatchedModelView {
  const PlanType &plan;
  std::size_t states;
  std::size_t batch;
  std::span<Scalar> node_potentials;
  std::span<Scalar> edge_potentials;
  operator BatchedModelView() const { return {plan, states, batch, node_potentials, edge_potentials}; }
};

struct BatchedCategoricalModelView {
  const PlanType &plan;
  std::size_t states;
  std::size_t batch;
  std::size_t categories;
  std::span<const tree_hmm_amd::Index> observation_nodes;
  std::span<const std::uint8_t> observations;
  std::span<const Scalar> root_potential;
  std::span<const Scalar> emission_potentials;
  std::span<const Scalar> edge_potentials;
};

struct MutableBatchedCategoricalModelView {
  const PlanType &plan;
  std::size_t states;
  std::size_t batch;
  std::size_t categories;
  std::span<const tree_hmm_amd::Index> observation_nodes;
  std::span<std::uint8_t> observations;
  std::span<Scalar> root_potential;
  std::span<Scalar> emission_potentials;
  std::span<Scalar> edge_potentials;
  operator BatchedCategoricalModelView() const {
    return {plan, states, batch, categories, observation_nodes, observations, root_potential, emission_potentials,
            edge_potentials};
  }
};

enum class CategoricalInputUpdate { kAll, kFactors, kNone };

struct AcceleratorTimings {
  double upload_ms = 0.0;
  double kernel_ms = 0.0;
  double download_ms = 0.0;
  double wall_ms = 0.0;
};

struct PartitionView {
  std::span<const Scalar> values;
  AcceleratorTimings timings;
};

struct BatchedMaximumAssignmentView {
  std::span<const Scalar> log_weights;
  std::span<const std::uint32_t> states;
  AcceleratorTimings timings;
};

struct BatchedPosteriorSampleView {
  std::span<const std::uint32_t> states;
  AcceleratorTimings timings;
};

struct BatchedMarginalView {
  std::span<const Scalar> log_partitions;
  std::span<const Scalar> nodes;
  std::span<const Scalar> edges;
  AcceleratorTimings timings;
};

namespace rocm {

inline bool Available() { return thmm_available() != 0; }

inline std::string DeviceDescription(int device = 0) {
  char buffer[256];
  tree_hmm_amd::Check(thmm_device_description(device, buffer, sizeof buffer));
  return buffer;
}

// include/tree_hmm/detail/gpu_backend_api.inc:4-95
class Workspace {
 public:
  Workspace() {
    thmm_workspace *raw = nullptr;
    tree_hmm_amd::Check(thmm_workspace_create(&raw));
    handle_.reset(raw);
  }
  Workspace(Workspace &&) noexcept = default;
  Workspace &operator=(Workspace &&) noexcept = default;
  Workspace(const Workspace &) = delete;
  Workspace &operator=(const Workspace &) = delete;

  void Reserve(const PlanType &plan, std::size_t states, std::size_t batch_capacity, int device = 0) {
    DoReserve(plan, states, batch_capacity, 0, device);
  }
  void ReserveMaximum(const PlanType &plan, std::size_t states, std::size_t batch_capacity, int device = 0) {
    DoReserve(plan, states, batch_capacity, THMM_RESERVE_MAXIMUM, device);
  }
  void ReserveSampling(const PlanType &plan, std::size_t states, std::size_t batch_capacity, int device = 0) {
    DoReserve(plan, states, batch_capacity, THMM_RESERVE_SAMPLING, device);
  }
  void ReserveMarginals(const PlanType &plan, std::size_t states, std::size_t batch_capacity, int device = 0) {
    DoReserve(plan, states, batch_capacity, THMM_RESERVE_MARGINALS, device);
  }
  void ReserveCategorical(const PlanType &plan, std::size_t states, std::size_t batch_capacity,
                          std::size_t categories, std::span<const tree_hmm_amd::Index> observation_nodes,
                          int device = 0) {
    DoReserveCategorical(plan, states, batch_capacity, categories, observation_nodes, 0, device);
  }
  void ReserveCategoricalMaximum(const PlanType &plan, std::size_t states, std::size_t batch_capacity,
                                 std::size_t categories, std::span<const tree_hmm_amd::Index> observation_nodes,
                                 int device = 0) {
    DoReserveCategorical(plan, states, batch_capacity, categories, observation_nodes, THMM_RESERVE_MAXIMUM, device);
  }
  void ReserveCategoricalSampling(const PlanType &plan, std::size_t states, std::size_t batch_capacity,
                                  std::size_t categories, std::span<const tree_hmm_amd::Index> observation_nodes,
                                  int device = 0) {
    DoReserveCategorical(plan, states, batch_capacity, categories, observation_nodes, THMM_RESERVE_SAMPLING, device);
  }
  void ReserveCategoricalMarginals(const PlanType &plan, std::size_t states, std::size_t batch_capacity,
                                   std::size_t categories, std::span<const tree_hmm_amd::Index> observation_nodes,
                                   int device = 0) {
    DoReserveCategorical(plan, states, batch_capacity, categories, observation_nodes, THMM_RESERVE_MARGINALS,
                         device);
  }

  MutableBatchedModelView Inputs() { return Inputs(capacity_); }
  MutableBatchedModelView Inputs(std::size_t batch) {
    void *nodes = nullptr, *edges = nullptr;
    std::size_t node_count = 0, edge_count = 0;
    tree_hmm_amd::Check(thmm_inputs(handle_.get(), batch, &nodes, &node_count, &edges, &edge_count));
    return {*plan_, states_, batch, {static_cast<Scalar *>(nodes), node_count},
            {static_cast<Scalar *>(edges), edge_count}};
  }
  MutableBatchedCategoricalModelView CategoricalInputs() { return CategoricalInputs(capacity_); }
  MutableBatchedCategoricalModelView CategoricalInputs(std::size_t batch) {
    std::uint8_t *obs = nullptr;
    void *root = nullptr, *emissions = nullptr, *edges = nullptr;
    std::size_t obs_count = 0, emission_count = 0, edge_count = 0;
    tree_hmm_amd::Check(thmm_categorical_inputs(handle_.get(), batch, &obs, &obs_count, &root, &emissions,
                                                &emission_count, &edges, &edge_count));
    return {*plan_,
            states_,
            batch,
            categories_,
            observation_nodes_,
            {obs, obs_count},
            {static_cast<Scalar *>(root), states_},
            {static_cast<Scalar *>(emissions), emission_count},
            {static_cast<Scalar *>(edges), edge_count}};
  }
  std::span<Scalar> Uniforms() { return Uniforms(capacity_); }
  std::span<Scalar> Uniforms(std::size_t batch) {
    void *ptr = nullptr;
    std::size_t count = 0;
    tree_hmm_amd::Check(thmm_uniforms(handle_.get(), batch, &ptr, &count));
    return {static_cast<Scalar *>(ptr), count};
  }

  thmm_workspace *handle() const { return handle_.get(); }

 private:
  struct Deleter {
    void operator()(thmm_workspace *w) const { thmm_workspace_destroy(w); }
  };
  void DoReserve(const PlanType &plan, std::size_t states, std::size_t batch, unsigned flags, int device) {
    binding_ = tree_hmm_amd::PlanBinding<PlanType>::Bind(plan);
    tree_hmm_amd::Check(
        thmm_reserve(handle_.get(), binding_.handle.get(), &plan, states, batch, kAmdPrecision, flags, device));
    plan_ = &plan;
    states_ = states;
    capacity_ = batch;
    categories_ = 0;
    observation_nodes_.clear();
  }
  void DoReserveCategorical(const PlanType &plan, std::size_t states, std::size_t batch, std::size_t categories,
                            std::span<const tree_hmm_amd::Index> observation_nodes, unsigned flags, int device) {
    binding_ = tree_hmm_amd::PlanBinding<PlanType>::Bind(plan);
    tree_hmm_amd::Check(thmm_reserve_categorical(handle_.get(), binding_.handle.get(), &plan, states, batch,
                                                 categories, observation_nodes.data(), observation_nodes.size(),
                                                 kAmdPrecision, flags, device));
    plan_ = &plan;
    states_ = states;
    capacity_ = batch;
    categories_ = categories;
    observation_nodes_.assign(observation_nodes.begin(), observation_nodes.end());
  }

  std::unique_ptr<thmm_workspace, Deleter> handle_;
  tree_hmm_amd::PlanBinding<PlanType> binding_;
  const PlanType *plan_ = nullptr;
  std::size_t states_ = 0, capacity_ = 0, categories_ = 0;
  std::vector<tree_hmm_amd::Index> observation_nodes_;
};

namespace detail {

inline AcceleratorTimings Timings(const thmm_timings &t) {
  return {t.upload_ms, t.kernel_ms, t.download_ms, t.wall_ms};
}

inline thmm_dense_model Dense(const BatchedModelView &m) {
  return {&m.plan, m.states, m.batch, m.node_potentials.data(), m.node_potentials.size(), m.edge_potentials.data(),
          m.edge_potentials.size()};
}

inline thmm_categorical_model Categorical(const BatchedCategoricalModelView &m) {
  return {&m.plan,
          m.states,
          m.batch,
          m.categories,
          m.observation_nodes.data(),
          m.observation_nodes.size(),
          m.observations.data(),
          m.observations.size(),
          m.root_potential.data(),
          m.root_potential.size(),
          m.emission_potentials.data(),
          m.emission_potentials.size(),
          m.edge_potentials.data(),
          m.edge_potentials.size()};
}

inline thmm_categorical_update Update(CategoricalInputUpdate u) {
  return u == CategoricalInputUpdate::kAll ? THMM_UPDATE_ALL
         : u == CategoricalInputUpdate::kFactors ? THMM_UPDATE_FACTORS
                                                 : THMM_UPDATE_NONE;
}

inline std::size_t Roots(const PlanType &plan) {
  if constexpr (std::is_same_v<PlanType, tree_hmm_amd::Plan>)
    return plan.roots().size();
  else
    return 1;
}

}  // namespace detail

// include/tree_hmm/detail/gpu_backend_api.inc:97-142
inline PartitionView PartitionFunctionPrepared(BatchedModelView model, Workspace &ws) {
  const thmm_dense_model m = detail::Dense(model);
  const void *values = nullptr;
  std::size_t count = 0;
  thmm_timings t{};
  tree_hmm_amd::Check(thmm_partition(ws.handle(), &m, 0, &values, &count, &t));
  return {{static_cast<const Scalar *>(values), count}, detail::Timings(t)};
}

inline PartitionView LogPartitionFunctionPrepared(BatchedModelView model, Workspace &ws) {
  const thmm_dense_model m = detail::Dense(model);
  const void *values = nullptr;
  std::size_t count = 0;
  thmm_timings t{};
  tree_hmm_amd::Check(thmm_partition(ws.handle(), &m, 1, &values, &count, &t));
  return {{static_cast<const Scalar *>(values), count}, detail::Timings(t)};
}

inline PartitionView PartitionFunctionPrepared(BatchedCategoricalModelView model, Workspace &ws,
                                               CategoricalInputUpdate update = CategoricalInputUpdate::kAll) {
  const thmm_categorical_model m = detail::Categorical(model);
  const void *values = nullptr;
  std::size_t count = 0;
  thmm_timings t{};
  tree_hmm_amd::Check(thmm_partition_categorical(ws.handle(), &m, detail::Update(update), 0, &values, &count, &t));
  return {{static_cast<const Scalar *>(values), count}, detail::Timings(t)};
}

inline PartitionView LogPartitionFunctionPrepared(BatchedCategoricalModelView model, Workspace &ws,
                                                  CategoricalInputUpdate update = CategoricalInputUpdate::kAll) {
  const thmm_categorical_model m = detail::Categorical(model);
  const void *values = nullptr;
  std::size_t count = 0;
  thmm_timings t{};
  tree_hmm_amd::Check(thmm_partition_categorical(ws.handle(), &m, detail::Update(update), 1, &values, &count, &t));
  return {{static_cast<const Scalar *>(values), count}, detail::Timings(t)};
}

inline BatchedMaximumAssignmentView MaximumAPosterioriPrepared(BatchedModelView model, Workspace &ws) {
  const thmm_dense_model m = detail::Dense(model);
  const void *weights = nullptr;
  const std::uint32_t *states = nullptr;
  thmm_timings t{};
  tree_hmm_amd::Check(thmm_maximum(ws.handle(), &m, &weights, &states, &t));
  const std::size_t roots = detail::Roots(model.plan);
  return {{static_cast<const Scalar *>(weights), model.batch * roots},
          {states, model.batch * model.plan.num_nodes()},
          detail::Timings(t)};
}

inline BatchedMaximumAssignmentView MaximumAPosterioriPrepared(
    BatchedCategoricalModelView model, Workspace &ws, CategoricalInputUpdate update = CategoricalInputUpdate::kAll) {
  const thmm_categorical_model m = detail::Categorical(model);
  const void *weights = nullptr;
  const std::uint32_t *states = nullptr;
  thmm_timings t{};
  tree_hmm_amd::Check(thmm_maximum_categorical(ws.handle(), &m, detail::Update(update), &weights, &states, &t));
  const std::size_t roots = detail::Roots(model.plan);
  return {{static_cast<const Scalar *>(weights), model.batch * roots},
          {states, model.batch * model.plan.num_nodes()},
          detail::Timings(t)};
}

inline BatchedPosteriorSampleView PosteriorSamplePrepared(BatchedModelView model, std::span<const Scalar> uniforms,
                                                          Workspace &ws) {
  const thmm_dense_model m = detail::Dense(model);
  const std::uint32_t *states = nullptr;
  thmm_timings t{};
  tree_hmm_amd::Check(thmm_sample(ws.handle(), &m, uniforms.data(), uniforms.size(), &states, &t));
  return {{states, model.batch * model.plan.num_nodes()}, detail::Timings(t)};
}

inline BatchedPosteriorSampleView PosteriorSamplePrepared(
    BatchedCategoricalModelView model, std::span<const Scalar> uniforms, Workspace &ws,
    CategoricalInputUpdate update = CategoricalInputUpdate::kAll) {
  const thmm_categorical_model m = detail::Categorical(model);
  const std::uint32_t *states = nullptr;
  thmm_timings t{};
  tree_hmm_amd::Check(
      thmm_sample_categorical(ws.handle(), &m, uniforms.data(), uniforms.size(), detail::Update(update), &states, &t));
  return {{states, model.batch * model.plan.num_nodes()}, detail::Timings(t)};
}

inline BatchedMarginalView PosteriorMarginalsPrepared(BatchedModelView model, Workspace &ws) {
  const thmm_dense_model m = detail::Dense(model);
  const void *log_z = nullptr, *nodes = nullptr, *edges = nullptr;
  thmm_timings t{};
  tree_hmm_amd::Check(thmm_marginals(ws.handle(), &m, &log_z, &nodes, &edges, &t));
  const std::size_t k = model.states;
  return {{static_cast<const Scalar *>(log_z), model.batch * detail::Roots(model.plan)},
          {static_cast<const Scalar *>(nodes), model.batch * model.plan.num_nodes() * k},
          {static_cast<const Scalar *>(edges), model.batch * model.plan.num_edges() * k * k},
          detail::Timings(t)};
}

inline BatchedMarginalView PosteriorMarginalsPrepared(BatchedCategoricalModelView model, Workspace &ws,
                                                      CategoricalInputUpdate update = CategoricalInputUpdate::kAll) {
  const thmm_categorical_model m = detail::Categorical(model);
  const void *log_z = nullptr, *nodes = nullptr, *edges = nullptr;
  thmm_timings t{};
  tree_hmm_amd::Check(thmm_marginals_categorical(ws.handle(), &m, detail::Update(update), &log_z, &nodes, &edges, &t));
  const std::size_t k = model.states;
  return {{static_cast<const Scalar *>(log_z), model.batch * detail::Roots(model.plan)},
          {static_cast<const Scalar *>(nodes), model.batch * model.plan.num_nodes() * k},
          {static_cast<const Scalar *>(edges), model.batch * model.plan.num_edges() * k * k},
          detail::Timings(t)};
}

}  // namespace rocm
}  // namespace tree_hmm

#endif  // TREE_HMM_AMD_ROCM_H_
