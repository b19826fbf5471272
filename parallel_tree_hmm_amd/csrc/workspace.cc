// Host executor: see workspace.h.
#include "workspace.h"

#include "expm.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>

namespace thmm {
namespace {

using Clock = std::chrono::steady_clock;

// HIP failures become runtime_error("<op>: <hipGetErrorString>")
// (src/gpu_backend_impl.inc:45-50).
void Hip(hipError_t status, const char *what) {
  if (status != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
}

void CheckLaunches() {
  const char *error = LastLaunchError();
  if (error != nullptr && error[0] != '\0') throw std::runtime_error(error);
}

// src/gpu_backend_impl.inc:52-67
std::uint32_t CheckedU32(std::size_t value, const char *what) {
  if (value > std::numeric_limits<std::uint32_t>::max()) throw std::length_error(std::string(what) + " exceeds uint32_t");
  return static_cast<std::uint32_t>(value);
}

std::size_t Product(std::initializer_list<std::size_t> values, const char *what) {
  std::size_t result = 1;
  for (std::size_t v : values) {
    if (v != 0 && result > std::numeric_limits<std::size_t>::max() / v)
      throw std::length_error(std::string(what) + " overflows size_t");
    result *= v;
  }
  return result;
}

std::size_t RealSize(Real real) { return real == Real::kF32 ? sizeof(float) : sizeof(double); }

// Owning device allocation (at least one byte so that every pointer is
// valid, as src/gpu_backend_impl.inc:69-75 does).
struct DeviceBuffer {
  void *ptr = nullptr;
  std::size_t bytes = 0;
  bool owned = true;  // false: a view into memory another buffer owns
  void Allocate(std::size_t size, const char *what) {
    Release();
    Hip(hipMalloc(&ptr, std::max<std::size_t>(size, 1)), what);
    bytes = size;
  }
  void View(void *at, std::size_t size) noexcept {
    Release();
    ptr = at;
    bytes = size;
    owned = false;
  }
  void Release() noexcept {
    if (ptr != nullptr && owned) static_cast<void>(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    owned = true;
  }
  template <typename V> V *as() const { return static_cast<V *>(ptr); }
  ~DeviceBuffer() { Release(); }
};

struct PinnedBuffer {
  void *ptr = nullptr;
  std::size_t bytes = 0;
  void Allocate(std::size_t size, const char *what) {
    Release();
    Hip(hipHostMalloc(&ptr, std::max<std::size_t>(size, 1), hipHostMallocDefault), what);
    bytes = size;
  }
  void Release() noexcept {
    if (ptr != nullptr) static_cast<void>(hipHostFree(ptr));
    ptr = nullptr;
    bytes = 0;
  }
  template <typename V> V *as() const { return static_cast<V *>(ptr); }
  ~PinnedBuffer() { Release(); }
};

struct Level {
  std::uint32_t primitive, offset, count;
  // Grouped compression levels: K-vectors moved per batch item in the
  // forward and reverse pass (profiling only).
  std::uint32_t forward_work = 0, reverse_work = 0;
};

}  // namespace

struct Workspace::Impl {
  const Plan *plan = nullptr;
  const void *identity = nullptr;
  std::size_t states = 0, capacity = 0, categories = 0;
  Real real = Real::kF32;
  unsigned flags = 0;
  int device = 0;
  bool categorical = false;
  std::vector<std::uint32_t> observation_nodes;
  std::size_t resident_observation_batch = 0;
  bool resident_factors = false;
  // Which forward pass the contraction state holds for a following backward
  // call; any other call overwrites that state and resets it.
  enum class Forward { kNone, kLogPotentials, kCategorical };
  Forward forward = Forward::kNone;
  std::size_t last_batch = 0;

  hipStream_t stream = nullptr;
  hipEvent_t events[6] = {};
  std::vector<Level> levels;
  // Every rake's branch ends up absorbed into the rake's parent (through
  // combinations that only merge branches), so a branch adjoint equals the
  // parent's node adjoint: the reverse pass skips the combination and
  // absorption levels and rakes read node_adj[parent].
  bool rake_adjoint_from_parent = false;
  // Sum-product forward program with fused rake/combination/absorption
  // levels (empty when the plan or state count does not allow it).
  struct FusedStep {
    std::uint32_t kind, offset, count;  // kind: 0 chunks, 1 segments, 3 compressions, 6 grouped compressions
    std::uint32_t work;                 // rakes (kind 0), partials (kind 1), compressions (kind 3), K-vectors (6)
  };
  std::vector<FusedStep> fused;
  // The same program with every compression level plain (kind 3): the FP64
  // sampling forward, which has no grouped compressions.
  std::vector<FusedStep> fused_plain;
  // Raw leaves: round 0's rakes (the first round0_rakes of the plan, every
  // leaf an original leaf) read the caller's potentials, as do round 0's
  // absorptions for their parents, and the node initialisation writes only
  // the other nodes (init_nodes, init_count of them, ascending); 0: not
  // available (round 0 does not start with rakes). round0_gather: the fused
  // forward program's first step holds exactly those rakes.
  std::uint32_t round0_rakes = 0;
  bool round0_gather = false;
  DeviceBuffer init_nodes;
  std::uint32_t init_count = 0;
  DeviceBuffer gather_rakes, gather_chunks, gather_segments, gather_partials;
  // Grouped compressions (K = 16 FP32 with the fused forward): the level of
  // each plan batch that holds compressions (count 0: none rooted there),
  // and the reverse program, the plan's levels with those in place of the
  // compression levels. Empty when compressions are not grouped.
  DeviceBuffer chains;
  std::vector<Level> chain_levels;
  std::vector<Level> grouped_levels;
  // Max-product compression pieces (MAP reservations at K = 8, 16, 32): the
  // plan's levels with each compression level replaced by the pieces rooted
  // there (empty when not built). compress_depth 1 runs the plain levels.
  DeviceBuffer max_pieces;
  // MAP's direct absorptions (with the pieces program): per rake, whether
  // its branch is absorbed as it stands, and the other absorptions, which
  // the program's absorption levels index (rake_direct stays null when a
  // rake level could leave the team kernels).
  DeviceBuffer max_direct, max_absorbs;
  std::vector<Level> max_levels;
  // Log-potential calls: the edges that compressions read (the only ones
  // staged when the rake kernels read log edges in place), the largest rake
  // level, and the caller's log edges a held forward pass read in place.
  DeviceBuffer compress_edges;
  std::uint32_t compress_edge_count = 0, max_rake_level = 0;
  const void *forward_log_edges = nullptr;
  // The log edges of the held forward pass and how it read them (a tree
  // group's backward half stages them again to recompute its paths).
  const void *forward_log_source = nullptr;
  bool forward_in_place = false, forward_compress_in_place = false;
  // Instantiated hipGraphs of the device-resident entry points (reserve flag
  // THMM_RESERVE_GRAPH), one per entry point, keyed by its arguments.
  struct GraphKey {
    const void *args[6] = {};
    std::size_t batch = 0;
    int flag = 0;
    bool operator==(const GraphKey &o) const {
      return std::equal(args, args + 6, o.args) && batch == o.batch && flag == o.flag;
    }
  };
  struct GraphCache {
    GraphKey key;
    hipGraphExec_t exec = nullptr;
    void Release() noexcept {
      if (exec != nullptr) static_cast<void>(hipGraphExecDestroy(exec));
      exec = nullptr;
    }
  };
  GraphCache graphs[4];
  Tuning tuning;

  // Budget for path storage (paths, their row scales and path adjoints) in
  // bytes, 0 for none; set before a reservation and kept across them.
  std::size_t path_budget = 0;
  // A forest whose path storage exceeds the budget runs in groups of
  // consecutive trees. Each group is a workspace reserved on the group's own
  // plan (local node, edge and root numbering) whose node state and edge row
  // shifts are views into this workspace's buffers, and whose paths, path
  // adjoints, branches and staged edges are views into one arena that every
  // group reuses in turn.
  struct TreeGroup {
    std::unique_ptr<Plan> plan;
    std::unique_ptr<Impl> w;
    std::size_t first_node = 0, first_edge = 0, first_root = 0;
  };
  std::vector<TreeGroup> groups;
  DeviceBuffer arena;
  // This workspace is a group of another one: its state buffers are views.
  bool group_member = false;

  // Task queues of the persistent passes (Tuning::persistent), built with
  // the fused forward program when the kernels cover it (K = 16 FP32, batch
  // 1, compression groups of at most 4 slots); num_tasks 0 otherwise.
  struct PersistentProgram {
    DeviceBuffer tasks, need, need_offset;
    std::uint32_t num_tasks = 0, steps = 0, groups = 0, grid = 0;
    std::uint64_t wait_ticks = 0;
    void Release() noexcept {
      tasks.Release(), need.Release(), need_offset.Release();
      num_tasks = steps = groups = grid = 0;
      wait_ticks = 0;
    }
  };
  PersistentProgram persistent_forward, persistent_reverse;
  DeviceBuffer persistent_counters, persistent_failure;

  // topology
  DeviceBuffer rakes, combines, absorbs, compresses, roots, is_root, observation_index;
  // Categorical leaf fusion: the first cat_rakes rakes (their leaves original
  // leaves) form their leaf vectors from the observations in the rake
  // kernels of partition and marginal passes; cat_skip [N] marks those
  // leaves for the node initialisation to skip. cat_rakes = 0: no fusion.
  DeviceBuffer cat_skip;
  std::uint32_t cat_rakes = 0;
  // batch sums of log Z (thmm_categorical_log_likelihood)
  DeviceBuffer site_sums;
  PinnedBuffer host_site_sums;
  // inputs
  PinnedBuffer host_nodes, host_edges, host_uniforms, host_observations, host_root, host_emissions;
  // observations [B][O] as uploaded, observations_t [O][B] as the kernels read them
  DeviceBuffer input_nodes, uniforms, observations, observations_t, root_potential, emissions;
  // scaled linear state (Scalar)
  DeviceBuffer edges, edge_scale, nodes, node_scale, paths, path_scale, branches, branch_scale, log_z;
  // adjoint
  DeviceBuffer node_adj, branch_adj, path_adj, node_out, edge_out;
  // FP64 scaled state for sampling when Scalar is float
  DeviceBuffer s_edges, s_nodes, s_node_scale, s_paths, s_path_scale, s_branches, s_branch_scale;
  DeviceBuffer assignment, failure;
  // max-product
  DeviceBuffer m_nodes, m_edges, m_paths, m_branches, rake_choice, compress_choice, log_weight;
  // results
  PinnedBuffer host_values, host_log_z, host_assignment, host_node_out, host_edge_out, host_failure;
  // branch-length chain (categorical marginal workspaces)
  DeviceBuffer chain_seed, chain_rate, chain_terms, chain_grad;
  PinnedBuffer host_chain_seed, host_chain_rate, host_chain_grad;
  // device expm of branch lengths: V, lambda, V_inv, then t (FP64)
  DeviceBuffer expm_factors;
  PinnedBuffer host_expm_factors;

  ~Impl() { Clear(); }

  void Clear() noexcept {
    if (stream != nullptr) {
      static_cast<void>(hipSetDevice(device));
      static_cast<void>(hipStreamSynchronize(stream));
      static_cast<void>(hipDeviceSynchronize());  // graph replays may run on other streams
    }
    for (GraphCache &g : graphs) g.Release();
    groups.clear();
    arena.Release();
    group_member = false;
    persistent_forward.Release();
    persistent_reverse.Release();
    persistent_counters.Release();
    persistent_failure.Release();
    tuning = Tuning{};
    for (DeviceBuffer *b : {&rakes, &combines, &absorbs, &compresses, &roots, &is_root, &observation_index,
                            &cat_skip, &site_sums, &init_nodes, &input_nodes, &uniforms, &observations,
                            &observations_t,
                            &root_potential, &emissions, &edges,
                            &edge_scale, &nodes, &node_scale, &paths, &path_scale, &branches, &branch_scale,
                            &log_z, &node_adj, &branch_adj, &path_adj, &node_out, &edge_out, &s_edges,
                            &s_nodes, &s_node_scale, &s_paths, &s_path_scale, &s_branches,
                            &s_branch_scale, &assignment, &failure, &m_nodes, &m_edges, &m_paths, &m_branches,
                            &rake_choice, &compress_choice, &log_weight, &chain_seed, &chain_rate, &chain_terms,
                            &chain_grad, &gather_rakes, &gather_chunks, &gather_segments, &gather_partials,
                            &chains, &max_pieces, &max_direct, &max_absorbs, &expm_factors, &compress_edges})
      b->Release();
    for (PinnedBuffer *b : {&host_nodes, &host_edges, &host_uniforms, &host_observations, &host_root,
                            &host_emissions, &host_values, &host_log_z, &host_assignment, &host_node_out,
                            &host_edge_out, &host_failure, &host_chain_seed, &host_chain_rate, &host_chain_grad,
                            &host_expm_factors, &host_site_sums})
      b->Release();
    for (hipEvent_t &e : events) {
      if (e != nullptr) static_cast<void>(hipEventDestroy(e));
      e = nullptr;
    }
    if (stream != nullptr) static_cast<void>(hipStreamDestroy(stream));
    stream = nullptr;
    plan = nullptr;
    identity = nullptr;
    states = capacity = categories = 0;
    flags = 0;
    device = 0;
    categorical = false;
    cat_rakes = 0;
    round0_rakes = 0;
    round0_gather = false;
    init_count = 0;
    observation_nodes.clear();
    resident_observation_batch = 0;
    resident_factors = false;
    last_batch = 0;
    forward = Forward::kNone;
    levels.clear();
    rake_adjoint_from_parent = false;
    fused.clear();
    fused_plain.clear();
    chain_levels.clear();
    grouped_levels.clear();
    max_levels.clear();
    compress_edge_count = max_rake_level = 0;
    forward_log_edges = nullptr;
    forward_log_source = nullptr;
    forward_in_place = forward_compress_in_place = false;
  }

  bool device_resident() const { return (flags & kReserveDeviceResident) != 0; }
  bool half_edges() const { return (flags & kReserveHalfEdgeMarginals) != 0; }
  std::size_t N() const { return plan->num_nodes; }
  std::size_t E() const { return plan->num_edges; }
  std::size_t R() const { return plan->num_branches; }
  std::size_t C() const { return plan->num_compressions(); }
  std::size_t Roots() const { return plan->roots.size(); }
  std::size_t KK() const { return states * states; }
  std::size_t S() const { return RealSize(real); }

  // The nodes a raw-leaf call initialises.
  NodeList InitList() const { return {init_nodes.as<std::uint32_t>(), init_count}; }
  GatherOps Gather() const {
    return {gather_rakes.as<DeviceRake>(), gather_chunks.as<DeviceGatherChunk>(),
            gather_segments.as<DeviceGatherSegment>(), gather_partials.as<double>()};
  }

  Launch MakeLaunch(Real r, std::size_t batch, hipStream_t s) const {
    Launch l{};
    l.real = r;
    l.dims = {static_cast<std::uint32_t>(states), static_cast<std::uint32_t>(batch), static_cast<std::uint32_t>(E()),
              static_cast<std::uint32_t>(N()), static_cast<std::uint32_t>(Roots())};
    l.ops = {rakes.as<DeviceRake>(), combines.as<DeviceCombine>(), absorbs.as<DeviceAbsorb>(),
             compresses.as<DeviceCompress>(), roots.as<std::uint32_t>(), chains.as<DeviceChain>(),
             max_pieces.as<DeviceMaxPiece>()};
    l.stream = s;
    l.tuning = tuning;
    return l;
  }

  // `edge_row_scales` is null for the caller's linear potentials and the
  // [E][K] row scales after a log-potential initialisation.
  LinearBuffers Linear(const void *edge_values, const double *edge_row_scales) const {
    return {nodes.ptr, node_scale.as<double>(), edge_values, edge_row_scales, paths.ptr,
            path_scale.as<double>(), branches.ptr, branch_scale.as<double>(), node_adj.ptr, branch_adj.ptr,
            path_adj.ptr, edge_out.ptr, nullptr, nullptr, half_edges() ? 1 : 0};
  }

  // `fp64_edges`: the edge values of an FP64 workspace's sampling call.
  LinearBuffers SamplingLinear(const void *fp64_edges = nullptr) const {
    if (real == Real::kF64) return Linear(fp64_edges != nullptr ? fp64_edges : edges.ptr, nullptr);
    return {s_nodes.ptr, s_node_scale.as<double>(), s_edges.ptr, nullptr, s_paths.ptr, s_path_scale.as<double>(),
            s_branches.ptr, s_branch_scale.as<double>(), nullptr, nullptr, nullptr, nullptr};
  }

  MaxBuffers Max() const {
    return {m_nodes.as<double>(), m_edges.as<double>(), m_paths.as<double>(), m_branches.as<double>(),
            rake_choice.as<std::uint8_t>(), compress_choice.as<std::uint8_t>(), nullptr, 0, nullptr, 0, nullptr};
  }
};

namespace {

using Impl = Workspace::Impl;

template <typename V> void Upload(DeviceBuffer &dst, const std::vector<V> &src, hipStream_t stream, const char *what) {
  dst.Allocate(src.size() * sizeof(V), what);
  if (!src.empty()) Hip(hipMemcpyAsync(dst.ptr, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice, stream), what);
}

// Partials per team of a partial-reduction level (the rakes per chunk of the
// fused forward, which a team walks serially, are w.tuning.gather_chunk).
constexpr std::uint32_t kGatherFanIn = 32;

// Groups the plan's batches into rounds of non-decreasing primitive kind
// (rakes, combinations, absorptions, compressions) and, when every rake's
// branch is absorbed in its own round, lays out the fused forward program:
// per round one chunk level over the rakes sorted by absorption, then for
// parents with more than one chunk of raked children levels of partial sums
// of fan-in kGatherFanIn until one sum per parent is absorbed, then the
// round's compressions. Leaves w.fused empty otherwise.
void BuildFusedForward(Impl &w, const Plan &plan, const std::vector<DeviceRake> &device_rakes, std::size_t batch) {
  std::vector<std::uint32_t> branch_absorb(plan.num_branches, 0xFFFFFFFFu);
  for (std::size_t a = 0; a < plan.absorptions.size(); ++a)
    branch_absorb[plan.absorptions[a].branch] = static_cast<std::uint32_t>(a);
  for (std::size_t i = plan.batches.size(); i-- > 0;) {
    const thmm_primitive_batch &b = plan.batches[i];
    if (b.primitive != THMM_BRANCH_COMBINATION) continue;
    for (std::uint32_t j = b.offset + b.count; j-- > b.offset;)
      branch_absorb[plan.combinations[j].source] = branch_absorb[plan.combinations[j].destination];
  }
  std::vector<std::uint32_t> absorb_round(plan.absorptions.size(), 0xFFFFFFFFu);
  std::vector<std::pair<std::size_t, std::size_t>> rounds;  // batch ranges
  for (std::size_t i = 0; i < plan.batches.size(); ++i) {
    if (i == 0 || plan.batches[i].primitive < plan.batches[i - 1].primitive) rounds.push_back({i, i});
    rounds.back().second = i + 1;
  }
  for (std::size_t r = 0; r < rounds.size(); ++r)
    for (std::size_t i = rounds[r].first; i < rounds[r].second; ++i)
      if (plan.batches[i].primitive == THMM_BRANCH_ABSORPTION)
        for (std::uint32_t j = plan.batches[i].offset; j < plan.batches[i].offset + plan.batches[i].count; ++j)
          absorb_round[j] = static_cast<std::uint32_t>(r);

  const std::uint32_t chunk_size = w.tuning.gather_chunk;
  std::vector<Impl::FusedStep> steps, plain;
  std::vector<DeviceRake> sorted;
  std::vector<DeviceGatherChunk> chunks;
  std::vector<DeviceGatherSegment> segments;
  std::uint32_t partials = 0;
  for (std::size_t r = 0; r < rounds.size(); ++r) {
    std::vector<std::pair<std::uint32_t, std::uint32_t>> keyed;  // (absorption, rake)
    for (std::size_t i = rounds[r].first; i < rounds[r].second; ++i) {
      const thmm_primitive_batch &b = plan.batches[i];
      if (b.primitive != THMM_RAKE) continue;
      for (std::uint32_t j = b.offset; j < b.offset + b.count; ++j) {
        const std::uint32_t a = branch_absorb[plan.rakes[j].branch];
        if (a == 0xFFFFFFFFu || absorb_round[a] != r || plan.absorptions[a].parent != plan.rakes[j].parent) return;
        keyed.push_back({a, j});
      }
    }
    std::stable_sort(keyed.begin(), keyed.end(),
                     [](const auto &x, const auto &y) { return x.first < y.first; });
    const std::uint32_t chunk_offset = static_cast<std::uint32_t>(chunks.size());
    // Parents whose chunks leave partials: (parent, first partial, count).
    std::vector<std::array<std::uint32_t, 3>> open;
    for (std::size_t i = 0; i < keyed.size();) {
      std::size_t end = i;
      while (end < keyed.size() && keyed[end].first == keyed[i].first) ++end;
      const std::uint32_t parent = plan.absorptions[keyed[i].first].parent;
      const bool direct = end - i <= chunk_size;
      const std::uint32_t first_partial = partials;
      for (std::size_t c = i; c < end; c += chunk_size) {
        const std::uint32_t count = static_cast<std::uint32_t>(std::min<std::size_t>(chunk_size, end - c));
        const DeviceRake &head = device_rakes[keyed[c].second];
        chunks.push_back({static_cast<std::uint32_t>(sorted.size()), count, parent, direct ? kDirectChunk : partials++,
                          head.leaf, head.slot, 0, 0});
        for (std::size_t k = c; k < c + count; ++k) sorted.push_back(device_rakes[keyed[k].second]);
      }
      if (!direct) open.push_back({parent, first_partial, partials - first_partial});
      i = end;
    }
    if (chunks.size() > chunk_offset) {
      steps.push_back({0, chunk_offset, static_cast<std::uint32_t>(chunks.size() - chunk_offset),
                       static_cast<std::uint32_t>(keyed.size())});
      plain.push_back(steps.back());
    }
    // Reduction levels: runs of up to kGatherFanIn partials per team; a
    // parent whose partials fit one run is absorbed at that level.
    while (!open.empty()) {
      const std::uint32_t segment_offset = static_cast<std::uint32_t>(segments.size());
      std::uint32_t parts = 0;
      std::vector<std::array<std::uint32_t, 3>> next;
      for (const auto &[parent, first, count] : open) {
        parts += count;
        if (count <= kGatherFanIn) {
          segments.push_back({parent, first, count, kDirectChunk});
          continue;
        }
        const std::uint32_t next_first = partials;
        for (std::uint32_t c = 0; c < count; c += kGatherFanIn)
          segments.push_back({parent, first + c, std::min(kGatherFanIn, count - c), partials++});
        next.push_back({parent, next_first, partials - next_first});
      }
      steps.push_back({1, segment_offset, static_cast<std::uint32_t>(segments.size() - segment_offset), parts});
      plain.push_back(steps.back());
      open = std::move(next);
    }
    for (std::size_t i = rounds[r].first; i < rounds[r].second; ++i) {
      if (plan.batches[i].primitive != THMM_COMPRESSION) continue;
      plain.push_back({3, plan.batches[i].offset, plan.batches[i].count, plan.batches[i].count});
      if (w.chain_levels.empty()) {
        steps.push_back({3, plan.batches[i].offset, plan.batches[i].count, plan.batches[i].count});
      } else if (w.chain_levels[i].count != 0) {
        const Level &level = w.chain_levels[i];
        steps.push_back({kChainPrimitive, level.offset, level.count, level.forward_work});
      }
    }
  }
  if (sorted.size() != plan.rakes.size()) return;
  // The fused kernels index (item, lane) in 32 bits.
  for (const Impl::FusedStep &step : steps)
    if (static_cast<std::uint64_t>(step.count) * batch * 64 >= (1ull << 32)) return;
  Upload(w.gather_rakes, sorted, w.stream, "fused topology upload");
  Upload(w.gather_chunks, chunks, w.stream, "fused topology upload");
  Upload(w.gather_segments, segments, w.stream, "fused topology upload");
  w.gather_partials.Allocate(Product({partials, batch, w.states, sizeof(double)}, "partial messages"),
                             "partial messages");
  // The uploads read the local vectors above.
  Hip(hipStreamSynchronize(w.stream), "fused topology upload");
  w.fused = std::move(steps);
  w.fused_plain = std::move(plain);
}

// Groups the renamed compressions for the K = 16 FP32 MFMA path. A
// compression's height is 1 plus the largest height of the compressions
// producing its inputs. Groups are formed top-down (consumers come after
// their inputs in plan order): a compression whose slot feeds a rake or a
// root starts a group; a group takes in an input compression, in place and
// left first, when that compression's height is not a multiple of
// w.tuning.compress_depth and the group stays within 2^depth slots (<=
// kMaxChain); every other input compression starts a group of its own.
// Balanced compression trees thus split into depth-level groups. A group
// runs at its root's level: every input is an edge or an earlier group's
// slot, and every middle node's value is final by then.
void BuildChains(Impl &w, const Plan &plan, const std::vector<DeviceCompress> &compresses, std::size_t batch) {
  w.chain_levels.clear();
  w.grouped_levels.clear();
  w.chains.Release();
  const std::uint32_t depth = w.tuning.compress_depth;
  if (depth < 2 || w.states != 16 || w.real != Real::kF32 || compresses.empty()) return;
  const std::size_t limit = std::min<std::size_t>(std::size_t{1} << depth, kMaxChain);
  const std::uint32_t E = static_cast<std::uint32_t>(plan.num_edges);
  std::vector<std::uint32_t> height(compresses.size(), 0);
  std::vector<std::uint8_t> root(compresses.size(), 1);  // starts a group
  for (const thmm_primitive_batch &b : plan.batches) {
    if (b.primitive != THMM_COMPRESSION) continue;
    for (std::uint32_t c = b.offset; c < b.offset + b.count; ++c) {
      std::uint32_t h = 1;
      for (std::uint32_t slot : {compresses[c].left, compresses[c].right})
        if (slot >= E) h = std::max(h, height[slot - E] + 1), root[slot - E] = 0;
      height[c] = h;
    }
  }
  // In-order inputs and middles of the group rooted at c. With `decide`,
  // input compressions that the rule above leaves out become group roots;
  // without, every input compression that is not a root is taken in.
  std::vector<std::uint32_t> slots, middles;
  const auto expand = [&](std::uint32_t c, bool decide) {
    slots = {compresses[c].left, compresses[c].right};
    middles = {compresses[c].middle};
    for (std::size_t k = 0; k < slots.size();) {
      const std::uint32_t slot = slots[k];
      if (slot >= E && !root[slot - E]) {
        if (!decide || (height[slot - E] % depth != 0 && slots.size() < limit)) {
          const DeviceCompress &op = compresses[slot - E];
          slots[k] = op.left;
          slots.insert(slots.begin() + static_cast<std::ptrdiff_t>(k) + 1, op.right);
          middles.insert(middles.begin() + static_cast<std::ptrdiff_t>(k), op.middle);
          continue;  // look at the new left input
        }
        root[slot - E] = 1;
      }
      ++k;
    }
  };
  for (std::size_t i = plan.batches.size(); i-- > 0;) {
    const thmm_primitive_batch &b = plan.batches[i];
    if (b.primitive != THMM_COMPRESSION) continue;
    for (std::uint32_t c = b.offset + b.count; c-- > b.offset;)
      if (root[c]) expand(c, true);
  }
  std::vector<DeviceChain> chains;
  std::vector<Level> levels(plan.batches.size(), Level{kChainPrimitive, 0, 0});
  // Per compression batch, groups of 2 slots first, then those of 3 or 4:
  // the reverse pass launches each range separately (kPairChainPrimitive,
  // kSmallChainPrimitive: fewer registers, more resident wavefronts).
  std::vector<Level> pairs(plan.batches.size(), Level{kPairChainPrimitive, 0, 0});
  std::vector<Level> small(plan.batches.size(), Level{kSmallChainPrimitive, 0, 0});
  const auto forward_work = [](std::uint32_t n) { return (n + 1) * 16 + (n - 1); };
  const auto reverse_work = [](std::uint32_t n) { return (2 * n + 1) * 16 + 2 * (n - 1); };
  for (std::size_t i = 0; i < plan.batches.size(); ++i) {
    const thmm_primitive_batch &b = plan.batches[i];
    if (b.primitive != THMM_COMPRESSION) continue;
    Level &level = levels[i];
    level.offset = static_cast<std::uint32_t>(chains.size());
    for (std::uint32_t c = b.offset; c < b.offset + b.count; ++c) {
      if (!root[c]) continue;
      expand(c, false);
      if (slots.size() < 2 || slots.size() > limit) throw std::logic_error("compression group exceeds its input limit");
      DeviceChain group{};
      group.count = static_cast<std::uint32_t>(slots.size());
      group.out = compresses[c].out;
      std::copy(slots.begin(), slots.end(), group.slots);
      std::copy(middles.begin(), middles.end(), group.middles);
      chains.push_back(group);
      level.forward_work += forward_work(group.count);
      level.reverse_work += reverse_work(group.count);
    }
    level.count = static_cast<std::uint32_t>(chains.size()) - level.offset;
    const auto first = chains.begin() + level.offset;
    const auto split =
        std::stable_partition(first, chains.end(), [](const DeviceChain &x) { return x.count <= 4; });
    const auto two = std::stable_partition(first, split, [](const DeviceChain &x) { return x.count == 2; });
    pairs[i].offset = level.offset;
    pairs[i].count = static_cast<std::uint32_t>(two - first);
    small[i].offset = pairs[i].offset + pairs[i].count;
    small[i].count = static_cast<std::uint32_t>(split - two);
    for (auto it = first; it != split; ++it) {
      Level &part = it < two ? pairs[i] : small[i];
      part.forward_work += forward_work(it->count);
      part.reverse_work += reverse_work(it->count);
    }
    // The kernels index (group, item) in 32 bits.
    if (static_cast<std::uint64_t>(level.count) * batch * 64 >= (1ull << 32)) return;
  }
  Upload(w.chains, chains, w.stream, "grouped compression upload");
  Hip(hipStreamSynchronize(w.stream), "grouped compression upload");
  for (std::size_t i = 0; i < plan.batches.size(); ++i) {
    if (plan.batches[i].primitive != THMM_COMPRESSION) {
      w.grouped_levels.push_back(w.levels[i]);
      continue;
    }
    const Level &all = levels[i];
    if (pairs[i].count != 0) w.grouped_levels.push_back(pairs[i]);
    if (small[i].count != 0) w.grouped_levels.push_back(small[i]);
    const std::uint32_t held = pairs[i].count + small[i].count;
    if (all.count > held)
      w.grouped_levels.push_back({kChainPrimitive, all.offset + held, all.count - held,
                                  all.forward_work - small[i].forward_work - pairs[i].forward_work,
                                  all.reverse_work - small[i].reverse_work - pairs[i].reverse_work});
  }
  w.chain_levels = std::move(levels);
}

// The fused forward program and, for K = 16 FP32, the grouped compressions
// it runs; grouping is dropped when the fused program cannot be built (the
// plan-level forward stores every compression).
template <typename V> std::vector<V> Download(const DeviceBuffer &src, std::size_t count, const char *what) {
  std::vector<V> out(count);
  if (count != 0) Hip(hipMemcpy(out.data(), src.ptr, count * sizeof(V), hipMemcpyDeviceToHost), what);
  return out;
}

// Tree of every node (its root's position in plan.roots).
std::vector<std::uint32_t> TreeOfNodes(const Plan &plan) {
  std::vector<std::vector<std::uint32_t>> children(plan.num_nodes);
  for (std::size_t e = 0; e < plan.num_edges; ++e) children[plan.edge_parents[e]].push_back(plan.edge_children[e]);
  std::vector<std::uint32_t> tree(plan.num_nodes, 0), stack;
  for (std::size_t r = 0; r < plan.roots.size(); ++r) {
    stack.push_back(plan.roots[r]);
    while (!stack.empty()) {
      const std::uint32_t v = stack.back();
      stack.pop_back();
      tree[v] = static_cast<std::uint32_t>(r);
      for (std::uint32_t c : children[v]) stack.push_back(c);
    }
  }
  return tree;
}

// The persistent passes are switched off: two runs on the MI355X (a 115 k-node
// K = 16 forest) did not complete within 120 and 60 seconds, the second with
// every dependency wait bounded at 2 s of wall clock, so the launches did not
// finish for a reason the waits do not explain. Their queues are not built
// and Tuning::persistent = 1 is refused until the cause is shown.
constexpr bool kPersistentPassesEnabled = false;

// Lays out the task queues of the persistent passes: every level of a pass
// (in execution order) is cut at tree-group boundaries and into tasks of at
// most kTaskBlocks blocks; a task's step is the index of its level among the
// levels in which its group has ops. Groups are consecutive trees of about
// equal node count, at most kPersistentGroups of them; when a level's ops
// are not ordered by group, all trees form one group (level barriers).
void BuildPersistent(Impl &w, std::size_t batch) {
  w.persistent_forward.Release();
  w.persistent_reverse.Release();
  if (!kPersistentPassesEnabled) return;
  if (w.states != 16 || w.real != Real::kF32 || batch != 1 || w.fused.empty() || w.chain_levels.empty()) return;
  for (const Impl::FusedStep &step : w.fused)
    if (step.kind != 0 && step.kind != 1 && step.kind != kChainPrimitive) return;
  for (const Level &level : w.grouped_levels)
    if (level.primitive == kChainPrimitive) return;  // groups of more than 4 slots
  constexpr std::uint32_t kPersistentGroups = 256, kTaskBlocks = 16;
  const Plan &plan = *w.plan;
  const std::vector<std::uint32_t> tree = TreeOfNodes(plan);
  std::vector<std::size_t> tree_nodes(plan.roots.size(), 0);
  for (std::uint32_t t : tree) ++tree_nodes[t];
  std::vector<std::uint32_t> group_of_tree(plan.roots.size(), 0);
  {
    const std::size_t target = (plan.num_nodes + kPersistentGroups - 1) / kPersistentGroups;
    std::uint32_t g = 0;
    std::size_t filled = 0;
    for (std::size_t t = 0; t < tree_nodes.size(); ++t) {
      if (filled >= target && g + 1 < kPersistentGroups) ++g, filled = 0;
      group_of_tree[t] = g;
      filled += tree_nodes[t];
    }
  }
  Hip(hipStreamSynchronize(w.stream), "persistent program download");
  const auto chunks = Download<DeviceGatherChunk>(w.gather_chunks, w.gather_chunks.bytes / sizeof(DeviceGatherChunk),
                                                  "persistent program download");
  const auto segments = Download<DeviceGatherSegment>(
      w.gather_segments, w.gather_segments.bytes / sizeof(DeviceGatherSegment), "persistent program download");
  const auto chains = Download<DeviceChain>(w.chains, w.chains.bytes / sizeof(DeviceChain), "persistent program download");
  const auto rakes = Download<DeviceRake>(w.rakes, plan.rakes.size(), "persistent program download");
  const std::uint32_t E = static_cast<std::uint32_t>(plan.num_edges);
  // One pass: (kind, ops per block, node whose tree owns op i) per level.
  struct PassLevel {
    std::uint32_t kind, offset, count, per_block;
    std::function<std::uint32_t(std::uint32_t)> node;
  };
  const auto chain_node = [&](std::uint32_t i) { return plan.compressions[chains[i].out - E].middle; };
  std::vector<PassLevel> forward, reverse;
  for (const Impl::FusedStep &step : w.fused) {
    if (step.kind == 0)
      forward.push_back({kTaskGather, step.offset, step.count, 16, [&](std::uint32_t i) { return chunks[i].parent; }});
    else if (step.kind == 1)
      forward.push_back({kTaskSegments, step.offset, step.count, 16, [&](std::uint32_t i) { return segments[i].parent; }});
    else
      forward.push_back({kTaskChainForward, step.offset, step.count, 4, chain_node});
  }
  for (std::size_t i = w.grouped_levels.size(); i-- > 0;) {
    const Level &level = w.grouped_levels[i];
    if (level.primitive == THMM_RAKE)
      reverse.push_back({kTaskRakeReverse, level.offset, level.count, 16, [&](std::uint32_t j) { return rakes[j].parent; }});
    else if (level.primitive == kPairChainPrimitive)
      reverse.push_back({kTaskChainReversePair, level.offset, level.count, 4, chain_node});
    else if (level.primitive == kSmallChainPrimitive)
      reverse.push_back({kTaskChainReverseSmall, level.offset, level.count, 4, chain_node});
  }
  const auto build = [&](const std::vector<PassLevel> &levels, Impl::PersistentProgram &out) {
    std::uint32_t groups = kPersistentGroups;
    for (const PassLevel &level : levels)
      for (std::uint32_t i = level.offset + 1; i < level.offset + level.count && groups > 1; ++i)
        if (group_of_tree[tree[level.node(i)]] < group_of_tree[tree[level.node(i - 1)]]) groups = 1;
    const auto group_of = [&](std::uint32_t node) { return groups == 1 ? 0u : group_of_tree[tree[node]]; };
    std::vector<DeviceTask> tasks;
    std::vector<std::vector<std::uint32_t>> need(kPersistentGroups);
    for (const PassLevel &level : levels) {
      const std::uint32_t cap = level.per_block * kTaskBlocks;
      for (std::uint32_t lo = level.offset; lo < level.offset + level.count;) {
        const std::uint32_t g = group_of(level.node(lo));
        std::uint32_t hi = lo + 1;
        while (hi < level.offset + level.count && group_of(level.node(hi)) == g) ++hi;
        const std::uint32_t step = static_cast<std::uint32_t>(need[g].size());
        need[g].push_back(0);
        for (std::uint32_t o = lo; o < hi; o += cap) {
          const std::uint32_t n = std::min(cap, hi - o);
          tasks.push_back({level.kind, step, g, o, n, (n + level.per_block - 1) / level.per_block, 0, 0});
          ++need[g][step];
        }
        lo = hi;
      }
    }
    std::vector<std::uint32_t> flat, offsets;
    for (const std::vector<std::uint32_t> &n : need) {
      offsets.push_back(static_cast<std::uint32_t>(flat.size()));
      flat.insert(flat.end(), n.begin(), n.end());
    }
    Upload(out.tasks, tasks, w.stream, "persistent program upload");
    Upload(out.need, flat, w.stream, "persistent program upload");
    Upload(out.need_offset, offsets, w.stream, "persistent program upload");
    out.num_tasks = static_cast<std::uint32_t>(tasks.size());
    out.steps = static_cast<std::uint32_t>(flat.size());
    out.groups = kPersistentGroups;
    Hip(hipStreamSynchronize(w.stream), "persistent program upload");
  };
  build(forward, w.persistent_forward);
  build(reverse, w.persistent_reverse);
  w.persistent_forward.grid = PersistentGrid(false);
  w.persistent_reverse.grid = PersistentGrid(true);
  // A dependency wait far longer than any level (the largest take a few
  // milliseconds) is reported as a failure rather than waited out.
  w.persistent_forward.wait_ticks = w.persistent_reverse.wait_ticks = WallClockTicks(2.0);
  const std::size_t words = 1 + std::max(w.persistent_forward.steps, w.persistent_reverse.steps) + kPersistentGroups;
  w.persistent_counters.Allocate(words * sizeof(std::uint32_t), "persistent pass counters");
  if (w.persistent_failure.ptr == nullptr) {
    w.persistent_failure.Allocate(sizeof(std::uint32_t), "persistent pass status");
    Hip(hipMemset(w.persistent_failure.ptr, 0, sizeof(std::uint32_t)), "persistent pass status");
  }
}

// Max-product compression pieces: a compression whose path only a later
// compression reads is absorbed into that consumer's piece unless one of
// its own inputs was absorbed (a greedy pass in plan order, children before
// parents), so a piece is a root and at most its two input compressions.
// On the 4,096-tree K = 16 forest this absorbs 53 % of the compressions.
void BuildMaxPieces(Impl &w, const Plan &plan, const std::vector<DeviceCompress> &compresses, std::size_t batch) {
  w.max_levels.clear();
  w.max_pieces.Release();
  w.max_direct.Release();
  w.max_absorbs.Release();
  const std::size_t K = w.states;
  if (!(w.flags & THMM_RESERVE_MAXIMUM) || (K != 8 && K != 16 && K != 32) || compresses.empty()) return;
  const std::uint32_t E = static_cast<std::uint32_t>(plan.num_edges);
  const std::uint32_t kNoConsumer = 0xFFFFFFFFu;
  std::vector<std::uint32_t> consumer(compresses.size(), kNoConsumer);
  for (std::size_t c = 0; c < compresses.size(); ++c)
    for (std::uint32_t slot : {compresses[c].left, compresses[c].right})
      if (slot >= E) consumer[slot - E] = static_cast<std::uint32_t>(c);
  std::vector<std::uint8_t> absorbed(compresses.size(), 0);
  for (const thmm_primitive_batch &b : plan.batches) {
    if (b.primitive != THMM_COMPRESSION) continue;
    for (std::uint32_t c = b.offset; c < b.offset + b.count; ++c) {
      bool take = consumer[c] != kNoConsumer;
      for (std::uint32_t slot : {compresses[c].left, compresses[c].right})
        if (slot >= E && absorbed[slot - E]) take = false;
      absorbed[c] = take ? 1 : 0;
    }
  }
  // Direct absorptions: a rake whose branch no combination touches is
  // absorbed by its own rake kernel; the absorption levels keep the others
  // (their ops in max_absorbs). Only where every rake level stays on the
  // team kernels.
  std::vector<std::uint8_t> combined(plan.num_branches, 0);
  for (const thmm_branch_combination &c : plan.combinations) combined[c.source] = combined[c.destination] = 1;
  bool direct = true;
  for (const thmm_primitive_batch &b : plan.batches)
    if (b.primitive == THMM_RAKE && static_cast<std::uint64_t>(b.count) * batch * K * K / 4 >= kMaxLaunchItems)
      direct = false;
  std::vector<DeviceAbsorb> absorbs;
  std::vector<DeviceMaxPiece> pieces;
  std::vector<Level> levels;
  for (const thmm_primitive_batch &b : plan.batches) {
    if (b.primitive == THMM_BRANCH_ABSORPTION && direct) {
      const std::uint32_t first = static_cast<std::uint32_t>(absorbs.size());
      for (std::uint32_t a = b.offset; a < b.offset + b.count; ++a)
        if (combined[plan.absorptions[a].branch])
          absorbs.push_back({static_cast<std::uint32_t>(plan.absorptions[a].parent), plan.absorptions[a].branch});
      const std::uint32_t count = static_cast<std::uint32_t>(absorbs.size()) - first;
      if (count != 0) levels.push_back({b.primitive, first, count});
      continue;
    }
    if (b.primitive != THMM_COMPRESSION) {
      levels.push_back({b.primitive, b.offset, b.count});
      continue;
    }
    const std::uint32_t first = static_cast<std::uint32_t>(pieces.size());
    for (std::uint32_t c = b.offset; c < b.offset + b.count; ++c) {
      if (absorbed[c]) continue;
      std::uint32_t flags = 0;
      if (compresses[c].left >= E && absorbed[compresses[c].left - E]) flags |= 1u;
      if (compresses[c].right >= E && absorbed[compresses[c].right - E]) flags |= 2u;
      pieces.push_back({c, flags});
    }
    const std::uint32_t count = static_cast<std::uint32_t>(pieces.size()) - first;
    // The team kernel covers a level in one dispatch below 2^31 work-items.
    if (static_cast<std::uint64_t>(count) * batch * K * K / 4 >= kMaxLaunchItems) return;
    if (count != 0) levels.push_back({kMaxPiecePrimitive, first, count});
  }
  Upload(w.max_pieces, pieces, w.stream, "max-product piece upload");
  if (direct) {
    std::vector<std::uint8_t> flags(plan.rakes.size());
    for (std::size_t r = 0; r < plan.rakes.size(); ++r) flags[r] = combined[plan.rakes[r].branch] ? 0 : 1;
    Upload(w.max_direct, flags, w.stream, "max-product direct absorption upload");
    Upload(w.max_absorbs, absorbs, w.stream, "max-product direct absorption upload");
  }
  Hip(hipStreamSynchronize(w.stream), "max-product piece upload");
  w.max_levels = std::move(levels);
}

void BuildPrograms(Impl &w, const Plan &plan, const std::vector<DeviceRake> &rakes,
                   const std::vector<DeviceCompress> &compresses, std::size_t batch) {
  w.fused.clear();
  w.fused_plain.clear();
  BuildChains(w, plan, compresses, batch);
  BuildFusedForward(w, plan, rakes, batch);
  if (w.fused.empty()) {
    w.chain_levels.clear();
    w.grouped_levels.clear();
    w.chains.Release();
  }
  BuildPersistent(w, batch);
}

void ReserveGroups(Impl &w, const Plan &plan, const void *identity, std::size_t states, Real real, unsigned flags,
                   int device, std::size_t slot_bytes);

// `group_member`: reserve a tree group of another workspace (ops and
// programs only; the owner assigns the state buffers as views).
void ReserveCommon(Impl &w, const Plan &plan, const void *identity, std::size_t states, std::size_t batch,
                   Real real, unsigned flags, int device, bool group_member = false) {
  if (states == 0 || batch == 0) throw std::invalid_argument("accelerator state and batch counts must be nonzero");
  if (states > 64) throw std::invalid_argument("state count exceeds the accelerator limit of 64");
  if (real != Real::kF32 && real != Real::kF64) throw std::invalid_argument("unknown precision");
  if ((flags & kReserveHalfEdgeMarginals) &&
      !((flags & kReserveDeviceResident) && (flags & THMM_RESERVE_MARGINALS)))
    throw std::invalid_argument("half-precision edge marginals need a device-resident marginals reservation");
  const std::size_t budget = w.path_budget;
  w.Clear();
  w.path_budget = budget;
  Hip(hipSetDevice(device), "GPU runtime SetDevice");
  w.plan = &plan;
  w.identity = identity;
  w.states = states;
  w.capacity = batch;
  w.real = real;
  w.flags = flags;
  w.device = device;
  CheckedU32(batch, "batch count");
  CheckedU32(Product({plan.num_nodes, batch}, "node-batch count"), "node-batch count");
  Hip(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking), "GPU runtime StreamCreateWithFlags");
  for (hipEvent_t &e : w.events) Hip(hipEventCreate(&e), "GPU runtime EventCreate");

  // Slot renaming: compression c writes the fresh slot E + c.
  const std::size_t E = plan.num_edges;
  RenamedOps renamed = RenameSlots(plan);
  const std::vector<DeviceRake> &rakes = renamed.rakes;
  const std::vector<DeviceCompress> &compresses = renamed.compresses;
  Upload(w.rakes, rakes, w.stream, "topology upload");
  Upload(w.combines, renamed.combines, w.stream, "topology upload");
  Upload(w.absorbs, renamed.absorbs, w.stream, "topology upload");
  Upload(w.compresses, compresses, w.stream, "topology upload");
  {
    std::vector<std::uint32_t> consumed;
    for (const DeviceCompress &c : compresses)
      for (std::uint32_t slot : {c.left, c.right})
        if (slot < E) consumed.push_back(slot);
    std::sort(consumed.begin(), consumed.end());
    Upload(w.compress_edges, consumed, w.stream, "topology upload");
    w.compress_edge_count = static_cast<std::uint32_t>(consumed.size());
    for (const thmm_primitive_batch &b : plan.batches)
      if (b.primitive == THMM_RAKE) w.max_rake_level = std::max(w.max_rake_level, b.count);
  }
  Upload(w.roots, plan.roots, w.stream, "topology upload");
  std::vector<std::uint8_t> is_root(plan.num_nodes, 0);
  for (std::uint32_t r : plan.roots) is_root[r] = 1;
  Upload(w.is_root, is_root, w.stream, "topology upload");
  for (const thmm_primitive_batch &b : plan.batches) w.levels.push_back({b.primitive, b.offset, b.count});
  {
    std::vector<std::uint32_t> branch_parent(plan.num_branches, 0xFFFFFFFFu);
    for (const thmm_branch_absorption &a : plan.absorptions) branch_parent[a.branch] = a.parent;
    for (std::size_t i = plan.batches.size(); i-- > 0;) {
      const thmm_primitive_batch &b = plan.batches[i];
      if (b.primitive != THMM_BRANCH_COMBINATION) continue;
      for (std::uint32_t j = b.offset + b.count; j-- > b.offset;)
        branch_parent[plan.combinations[j].source] = branch_parent[plan.combinations[j].destination];
    }
    w.rake_adjoint_from_parent = true;
    for (const thmm_rake &r : plan.rakes)
      if (branch_parent[r.branch] != r.parent) w.rake_adjoint_from_parent = false;
  }
  if (w.rake_adjoint_from_parent && FusedForwardSupported(static_cast<std::uint32_t>(states)))
    BuildPrograms(w, plan, rakes, compresses, batch);
  BuildMaxPieces(w, plan, compresses, batch);
  // Raw leaves where round 0 starts with a level of rakes of original leaves
  // (the per-call conditions are checked again per call).
  if (!plan.batches.empty() && plan.batches.front().primitive == THMM_RAKE && plan.batches.front().offset == 0) {
    const thmm_primitive_batch &b0 = plan.batches.front();
    std::vector<std::uint8_t> has_child(plan.num_nodes, 0), skip(plan.num_nodes, 0);
    for (Index parent : plan.edge_parents) has_child[parent] = 1;
    bool leaves = true;
    for (std::uint32_t r = 0; leaves && r < b0.count; ++r) {
      leaves = !has_child[plan.rakes[r].leaf];
      skip[plan.rakes[r].leaf] = 1;
    }
    // Round 0's absorbing parents are skipped too: with raw leaves the
    // fused program's round 0 (its gather and partial-sum levels) and MAP's
    // round-0 absorption levels form their vectors from the caller's
    // potentials when they first absorb into them, and nothing reads them
    // before. Each absorbs once in round 0 (else no raw leaves).
    for (std::size_t i = 0; leaves && i < plan.batches.size(); ++i) {
      const thmm_primitive_batch &b = plan.batches[i];
      if (i > 0 && b.primitive < plan.batches[i - 1].primitive) break;  // round 1 begins
      if (b.primitive == THMM_BRANCH_ABSORPTION)
        for (std::uint32_t j = b.offset; leaves && j < b.offset + b.count; ++j) {
          leaves = skip[plan.absorptions[j].parent] == 0;
          skip[plan.absorptions[j].parent] = 1;
        }
    }
    if (leaves) {
      std::vector<std::uint32_t> init;
      for (std::size_t n = 0; n < plan.num_nodes; ++n)
        if (skip[n] == 0) init.push_back(static_cast<std::uint32_t>(n));
      Upload(w.init_nodes, init, w.stream, "topology upload");
      w.init_count = static_cast<std::uint32_t>(init.size());
      w.round0_rakes = b0.count;
      w.round0_gather = !w.fused.empty() && w.fused.front().kind == 0 && w.fused.front().work == b0.count;
    }
  }

  const std::size_t K = states, KK = K * K, N = plan.num_nodes, R = plan.num_branches, C = plan.num_compressions();
  const std::size_t S = RealSize(real);
  const std::size_t roots = plan.roots.size();
  const bool resident = (flags & kReserveDeviceResident) != 0;
  const std::size_t node_values = Product({batch, N, K}, "node workspace");
  const std::size_t edge_values = Product({E, KK}, "edge inputs");
  const std::size_t path_values = Product({C, batch, KK}, "path workspace");
  const std::size_t branch_values = Product({R, batch, K}, "branch workspace");
  w.group_member = group_member;
  if (group_member) {
    Hip(hipStreamSynchronize(w.stream), "topology upload");
    return;
  }
  // Path storage per compression slot and item: the path, its row scales
  // and, for marginals, its adjoint.
  const std::size_t slot_bytes =
      KK * S + K * sizeof(double) + ((flags & THMM_RESERVE_MARGINALS) ? KK * S : 0);
  const std::size_t path_bytes = Product({C, batch, slot_bytes}, "path storage");
  if (w.path_budget != 0 && path_bytes > w.path_budget) {
    if (batch != 1 || !resident || (flags & (THMM_RESERVE_MAXIMUM | THMM_RESERVE_SAMPLING)) || roots < 2)
      throw std::length_error("path storage of " + std::to_string(path_bytes) + " bytes exceeds the budget of " +
                              std::to_string(w.path_budget) +
                              " bytes; only a device-resident batch-1 forest reservation without MAP or sampling "
                              "can run in tree groups");
    w.nodes.Allocate(node_values * S, "node storage");
    w.node_scale.Allocate(N * batch * sizeof(double), "node scales");
    w.edge_scale.Allocate(E * K * sizeof(double), "edge row scales");
    w.log_z.Allocate(batch * roots * sizeof(double), "log partitions");
    ReserveGroups(w, plan, identity, states, real, flags, device, slot_bytes);
    Hip(hipStreamSynchronize(w.stream), "topology upload");
    return;
  }

  // Device-resident reservations stage edges only for log-potential calls
  // that need it (EnsureEdgeStaging).
  if (!resident) w.edges.Allocate(edge_values * S, "edge storage");
  w.edge_scale.Allocate(E * K * sizeof(double), "edge row scales");
  w.nodes.Allocate(node_values * S, "node storage");
  w.node_scale.Allocate(N * batch * sizeof(double), "node scales");
  w.paths.Allocate(path_values * S, "path storage");
  w.path_scale.Allocate(C * batch * K * sizeof(double), "path row scales");
  w.branches.Allocate(branch_values * S, "branch storage");
  w.branch_scale.Allocate(R * batch * sizeof(double), "branch scales");
  w.log_z.Allocate(batch * roots * sizeof(double), "log partitions");
  if (!resident) {
    w.host_edges.Allocate(edge_values * S, "pinned edge inputs");
    w.host_values.Allocate(batch * roots * S, "pinned outputs");
    w.host_log_z.Allocate(batch * roots * sizeof(double), "pinned log partitions");
  }
  if (flags & THMM_RESERVE_MAXIMUM) {
    w.m_nodes.Allocate(node_values * sizeof(double), "MAP nodes");
    w.m_edges.Allocate(edge_values * sizeof(double), "MAP edges");
    w.m_paths.Allocate(path_values * sizeof(double), "MAP paths");
    w.m_branches.Allocate(branch_values * sizeof(double), "MAP branches");
    w.rake_choice.Allocate(branch_values, "MAP rake choices");
    w.compress_choice.Allocate(path_values, "MAP compression choices");
    w.log_weight.Allocate(batch * roots * sizeof(double), "MAP log weights");
    w.assignment.Allocate(N * batch * sizeof(std::uint32_t), "assignments");
    if (!resident) w.host_assignment.Allocate(N * batch * sizeof(std::uint32_t), "pinned assignments");
  }
  if (flags & THMM_RESERVE_SAMPLING) {
    if (real == Real::kF32) {
      w.s_edges.Allocate(edge_values * sizeof(double), "sampling edges");
      w.s_nodes.Allocate(node_values * sizeof(double), "sampling nodes");
      w.s_node_scale.Allocate(N * batch * sizeof(double), "sampling node scales");
      w.s_paths.Allocate(path_values * sizeof(double), "sampling paths");
      w.s_path_scale.Allocate(C * batch * K * sizeof(double), "sampling path row scales");
      w.s_branches.Allocate(branch_values * sizeof(double), "sampling branches");
      w.s_branch_scale.Allocate(R * batch * sizeof(double), "sampling branch scales");
    }
    w.uniforms.Allocate(N * batch * S, "posterior uniforms");
    w.failure.Allocate(sizeof(std::uint32_t), "sampling status");
    if (!w.assignment.ptr) w.assignment.Allocate(N * batch * sizeof(std::uint32_t), "assignments");
    if (!resident) {
      w.host_uniforms.Allocate(N * batch * S, "pinned uniforms");
      if (!w.host_assignment.ptr) w.host_assignment.Allocate(N * batch * sizeof(std::uint32_t), "pinned assignments");
      w.host_failure.Allocate(sizeof(std::uint32_t), "pinned sampling status");
    }
  }
  if (flags & THMM_RESERVE_MARGINALS) {
    w.node_adj.Allocate(node_values * S, "node adjoints");
    if (!w.rake_adjoint_from_parent) w.branch_adj.Allocate(branch_values * S, "branch adjoints");
    w.path_adj.Allocate(path_values * S, "path adjoints");
    if (!resident) {
      w.node_out.Allocate(node_values * S, "node marginals");
      w.edge_out.Allocate(Product({batch, E, KK}, "edge marginals") * S, "edge marginals");
      w.host_node_out.Allocate(node_values * S, "pinned node marginals");
      w.host_edge_out.Allocate(Product({batch, E, KK}, "edge marginals") * S, "pinned edge marginals");
    }
  }
  Hip(hipStreamSynchronize(w.stream), "topology upload");
}

// Splits a forest whose trees occupy consecutive node ranges into groups of
// consecutive trees whose path storage fits w.path_budget, reserves each
// group on its own plan and assigns its buffers (see Impl::TreeGroup). The
// group plans are planned from the groups' parent arrays: the planner
// contracts the trees of a forest independently, so each group plan holds
// the same ops, in the same order, as the forest plan holds for those trees
// (checked here through the edge arrays), and every kernel computes each op
// on its own, so grouped calls give the values of ungrouped ones bit for
// bit.
void ReserveGroups(Impl &w, const Plan &plan, const void *identity, std::size_t states, Real real, unsigned flags,
                   int device, std::size_t slot_bytes) {
  const std::size_t N = plan.num_nodes, E = plan.num_edges, T = plan.roots.size(), K = states, KK = K * K;
  const std::size_t S = RealSize(real);
  const auto too_big = [&](const std::string &why) {
    throw std::length_error("path storage exceeds the budget of " + std::to_string(w.path_budget) + " bytes and " +
                            why);
  };
  if (plan.roots.front() != 0) too_big("the forest does not start with a root");
  // Tree of every node: tree t spans [roots[t], roots[t + 1]).
  std::vector<std::uint32_t> tree(N);
  for (std::size_t t = 0; t < T; ++t) {
    const std::size_t end = t + 1 < T ? plan.roots[t + 1] : N;
    std::fill(tree.begin() + plan.roots[t], tree.begin() + static_cast<std::ptrdiff_t>(end), static_cast<std::uint32_t>(t));
  }
  std::vector<std::int64_t> parent(N, -1);
  for (std::size_t e = 0; e < E; ++e) {
    if (tree[plan.edge_parents[e]] != tree[plan.edge_children[e]])
      too_big("its trees do not occupy consecutive node ranges (tree groups need them)");
    parent[plan.edge_children[e]] = plan.edge_parents[e];
  }
  std::vector<std::size_t> per_tree(T, 0);
  for (const thmm_compression &c : plan.compressions) ++per_tree[tree[c.middle]];
  // Greedy groups of consecutive trees.
  std::vector<std::size_t> starts;
  std::size_t used = 0;
  for (std::size_t t = 0; t < T; ++t) {
    const std::size_t need = per_tree[t] * slot_bytes;
    if (need > w.path_budget)
      too_big("tree " + std::to_string(t) + " alone needs " + std::to_string(need) + " bytes of path storage");
    if (starts.empty() || used + need > w.path_budget) starts.push_back(t), used = 0;
    used += need;
  }
  starts.push_back(T);
  // Reserve the groups; the arena holds the largest group's pieces.
  const bool marginals = (flags & THMM_RESERVE_MARGINALS) != 0;
  const bool stage = !(real == Real::kF32 && (K == 32 || K == 64));
  std::size_t max_c = 0, max_r = 0, max_e = 0;
  bool branch_adjoints = false;
  for (std::size_t g = 0; g + 1 < starts.size(); ++g) {
    const std::size_t t0 = starts[g], t1 = starts[g + 1];
    const std::size_t n0 = plan.roots[t0], n1 = t1 < T ? plan.roots[t1] : N;
    std::vector<std::int64_t> local(n1 - n0);
    for (std::size_t i = n0; i < n1; ++i) local[i - n0] = parent[i] < 0 ? -1 : parent[i] - static_cast<std::int64_t>(n0);
    Impl::TreeGroup group;
    group.plan = std::make_unique<Plan>(MakePlan(local.data(), local.size()));
    group.first_node = n0;
    group.first_edge = n0 - t0;
    group.first_root = t0;
    const Plan &sub = *group.plan;
    bool same = sub.roots.size() == t1 - t0 && sub.num_edges == (n1 - n0) - (t1 - t0);
    for (std::size_t e = 0; same && e < sub.num_edges; ++e)
      same = plan.edge_children[group.first_edge + e] == n0 + sub.edge_children[e] &&
             plan.edge_parents[group.first_edge + e] == n0 + sub.edge_parents[e];
    if (!same) throw std::logic_error("a tree group's plan does not number its edges as the forest plan does");
    group.w = std::make_unique<Impl>();
    ReserveCommon(*group.w, sub, identity, states, 1, real, flags & ~kReserveGraph, device, true);
    max_c = std::max(max_c, sub.num_compressions());
    max_r = std::max(max_r, sub.num_branches);
    max_e = std::max(max_e, sub.num_edges);
    branch_adjoints = branch_adjoints || (marginals && !group.w->rake_adjoint_from_parent);
    w.groups.push_back(std::move(group));
  }
  const auto round = [](std::size_t bytes) { return (bytes + 255) & ~std::size_t{255}; };
  const std::size_t paths = round(max_c * KK * S), path_scale = round(max_c * K * sizeof(double));
  const std::size_t path_adj = marginals ? round(max_c * KK * S) : 0;
  const std::size_t branches = round(max_r * K * S), branch_scale = round(max_r * sizeof(double));
  const std::size_t branch_adj = branch_adjoints ? round(max_r * K * S) : 0;
  const std::size_t staging = stage ? round(max_e * KK * S) : 0;
  w.arena.Allocate(paths + path_scale + path_adj + branches + branch_scale + branch_adj + staging, "tree group arena");
  for (Impl::TreeGroup &group : w.groups) {
    Impl &gw = *group.w;
    char *at = static_cast<char *>(w.arena.ptr);
    const std::size_t n = gw.N(), e = gw.E();
    gw.nodes.View(static_cast<char *>(w.nodes.ptr) + group.first_node * K * S, n * K * S);
    gw.node_scale.View(w.node_scale.as<double>() + group.first_node, n * sizeof(double));
    gw.edge_scale.View(w.edge_scale.as<double>() + group.first_edge * K, e * K * sizeof(double));
    gw.paths.View(at, paths), at += paths;
    gw.path_scale.View(at, path_scale), at += path_scale;
    if (marginals) gw.path_adj.View(at, path_adj);
    at += path_adj;
    gw.branches.View(at, branches), at += branches;
    gw.branch_scale.View(at, branch_scale), at += branch_scale;
    if (marginals && !gw.rake_adjoint_from_parent) gw.branch_adj.View(at, branch_adj);
    at += branch_adj;
    if (stage) gw.edges.View(at, staging);
  }
}

void RequireHost(const Impl &w) {
  if (w.plan == nullptr) throw std::logic_error("accelerator Workspace::Reserve must precede this call");
  if (w.device_resident())
    throw std::logic_error("this workspace was reserved for device-resident calls only");
}

struct Timer {
  Clock::time_point wall = Clock::now();
};

void FillTimings(Impl &w, const Timer &timer, thmm_timings *t) {
  if (t == nullptr) return;
  float upload = 0.f, kernel = 0.f, download = 0.f;
  Hip(hipEventElapsedTime(&upload, w.events[0], w.events[1]), "GPU runtime EventElapsedTime upload");
  Hip(hipEventElapsedTime(&kernel, w.events[2], w.events[3]), "GPU runtime EventElapsedTime kernels");
  Hip(hipEventElapsedTime(&download, w.events[4], w.events[5]), "GPU runtime EventElapsedTime download");
  t->upload_ms = upload;
  t->kernel_ms = kernel;
  t->download_ms = download;
  t->wall_ms = std::chrono::duration<double, std::milli>(Clock::now() - timer.wall).count();
}

void Record(Impl &w, int index) { Hip(hipEventRecord(w.events[index], w.stream), "GPU runtime EventRecord"); }

// Validation shared by every prepared call (src/gpu_backend_impl.inc:2162-2173).
void CheckCall(const Impl &w, const void *identity, std::size_t states, std::size_t batch, std::size_t edge_count,
               const char *needs) {
  if (w.plan == nullptr || w.identity != identity || w.states != states || batch == 0 || batch > w.capacity)
    throw std::invalid_argument(std::string("prepared accelerator inference requires ") + needs +
                                " for this plan, state count, and batch capacity");
  if (edge_count != w.E() * w.KK()) throw std::invalid_argument("accelerator edge input shape is wrong");
}

// Stages the dense model into pinned memory (skipped for workspace-backed
// spans, src/gpu_backend_impl.inc:2820-2823) and enqueues the upload.
void StageDense(Impl &w, const thmm_dense_model &model) {
  if (w.categorical) throw std::invalid_argument("dense accelerator inference cannot use a categorical workspace");
  const std::size_t node_values = Product({model.batch, w.N(), w.states}, "node inputs");
  if (model.node_count != node_values) throw std::invalid_argument("accelerator node input shape is wrong");
  if (node_values != 0 && model.node_potentials == nullptr) throw std::invalid_argument("node potentials are null");
  if (model.edge_count != 0 && model.edge_potentials == nullptr) throw std::invalid_argument("edge potentials are null");
  if (model.node_potentials != w.host_nodes.ptr) std::memcpy(w.host_nodes.ptr, model.node_potentials, node_values * w.S());
  if (model.edge_potentials != w.host_edges.ptr && model.edge_count != 0)
    std::memcpy(w.host_edges.ptr, model.edge_potentials, model.edge_count * w.S());
}

void UploadDense(Impl &w, const thmm_dense_model &model) {
  Hip(hipMemcpyAsync(w.input_nodes.ptr, w.host_nodes.ptr, model.node_count * w.S(), hipMemcpyHostToDevice, w.stream),
      "GPU runtime MemcpyAsync node upload");
  if (model.edge_count != 0)
    Hip(hipMemcpyAsync(w.edges.ptr, w.host_edges.ptr, model.edge_count * w.S(), hipMemcpyHostToDevice, w.stream),
        "GPU runtime MemcpyAsync edge upload");
}


}  // namespace

// Event pairs around individual launches (measurement only).
class Profiler {
 public:
  Profiler(hipStream_t stream, std::uint32_t batch) : stream_(stream), batch_(batch) {}
  ~Profiler() {
    for (Entry &e : entries_) {
      static_cast<void>(hipEventDestroy(e.start));
      static_cast<void>(hipEventDestroy(e.stop));
    }
  }
  void Begin(std::uint32_t phase, std::uint32_t primitive, std::uint32_t count, std::uint32_t ops = ~0u) {
    Entry e{};
    e.info = {phase, primitive, count, batch_, 0.0, ops == ~0u ? count : ops};
    Hip(hipEventCreate(&e.start), "GPU runtime EventCreate");
    Hip(hipEventCreate(&e.stop), "GPU runtime EventCreate");
    Hip(hipEventRecord(e.start, stream_), "GPU runtime EventRecord");
    entries_.push_back(e);
  }
  void End() { Hip(hipEventRecord(entries_.back().stop, stream_), "GPU runtime EventRecord"); }
  std::size_t Report(thmm_launch_timing *out, std::size_t capacity) {
    Hip(hipStreamSynchronize(stream_), "profiled execution");
    for (std::size_t i = 0; i < entries_.size() && i < capacity; ++i) {
      float ms = 0.f;
      Hip(hipEventElapsedTime(&ms, entries_[i].start, entries_[i].stop), "GPU runtime EventElapsedTime");
      out[i] = entries_[i].info;
      out[i].ms = ms;
    }
    return entries_.size();
  }

 private:
  struct Entry {
    hipEvent_t start, stop;
    thmm_launch_timing info;
  };
  hipStream_t stream_;
  std::uint32_t batch_;
  std::vector<Entry> entries_;
};

namespace {

void ForwardLevels(const Launch &l, const std::vector<Level> &levels, const LinearBuffers &buf,
                   Profiler *profiler = nullptr) {
  for (const Level &level : levels) {
    if (profiler) profiler->Begin(1, level.primitive, level.count);
    LaunchForwardLevel(l, level.primitive, level.offset, level.count, buf);
    if (profiler) profiler->End();
  }
}

// Whether SumProductForward runs the fused program for these buffers (its
// kernels read edge slots with 16-byte accesses).
bool UsesFusedForward(const Impl &w, const LinearBuffers &buf) {
  return !w.fused.empty() && (reinterpret_cast<std::uintptr_t>(buf.edges) & 15u) == 0;
}

// The sum-product forward pass: the fused program when the workspace has
// one, else the plan's levels.
// `raw_leaves`: the caller's linear node potentials for round 0's gather
// (the first step) and its partial-sum levels (the steps of kind 1 that
// follow it), whose leaves and absorbing parents the node initialisation
// skipped.
void RunFusedSteps(const Impl &w, const std::vector<Impl::FusedStep> &steps, const Launch &l,
                   const LinearBuffers &buf, Profiler *profiler = nullptr, const void *raw_leaves = nullptr) {
  const GatherOps g = w.Gather();
  LinearBuffers raw = buf;
  raw.raw_leaves = raw_leaves;
  bool round0 = true;
  for (const Impl::FusedStep &step : steps) {
    round0 = round0 && (&step == &steps.front() || step.kind == 1);
    // Profiled as primitive 4 (fused rake chunks, count = rakes), 5 (partial
    // absorptions, count = partials), a compression level, or a grouped
    // compression level (kChainPrimitive, count = K-vectors moved per item).
    if (profiler)
      profiler->Begin(1, step.kind == 3 || step.kind == kChainPrimitive ? step.kind : 4 + step.kind, step.work,
                      step.count);
    if (step.kind == 0)
      LaunchGatherChunks(l, g, step.offset, step.count, round0 ? raw : buf);
    else if (step.kind == 1)
      LaunchGatherSegments(l, g, step.offset, step.count, round0 ? raw : buf);
    else
      LaunchForwardLevel(l, step.kind, step.offset, step.count, buf);
    if (profiler) profiler->End();
  }
}

// Profiled primitive of a persistent pass (count = its tasks).
constexpr std::uint32_t kPersistentPrimitive = 9;

// Whether this pass runs as one persistent launch: the tuning asks for it,
// the queues exist (K = 16 FP32, batch 1) and the buffers take the fused
// program with grouped compressions.
bool PersistentReady(const Impl &w, const Launch &l, const LinearBuffers &buf) {
  return w.tuning.persistent != 0 && w.persistent_forward.num_tasks != 0 && l.dims.B == 1 &&
         UsesFusedForward(w, buf) && buf.compress_log_edges == 0 && !w.grouped_levels.empty() &&
         buf.branch_adj == nullptr;
}

PersistentQueue QueueOf(const Impl &w, const Impl::PersistentProgram &p) {
  return {p.tasks.as<DeviceTask>(), p.num_tasks, p.steps, p.groups, p.need.as<std::uint32_t>(),
          p.need_offset.as<std::uint32_t>(), w.persistent_counters.as<std::uint32_t>(),
          w.persistent_failure.as<std::uint32_t>(), p.grid, p.wait_ticks};
}

// `raw_leaves`: see RawLeaves (only with the fused program).
void SumProductForward(const Impl &w, const Launch &l, const LinearBuffers &buf, Profiler *profiler = nullptr,
                       const void *raw_leaves = nullptr) {
  if (PersistentReady(w, l, buf)) {
    if (profiler) profiler->Begin(1, kPersistentPrimitive, w.persistent_forward.num_tasks);
    LaunchPersistentPass(l, false, w.Gather(), QueueOf(w, w.persistent_forward), buf);
    if (profiler) profiler->End();
  } else if (UsesFusedForward(w, buf)) {
    RunFusedSteps(w, w.fused, l, buf, profiler, raw_leaves);
  } else {
    ForwardLevels(l, w.levels, buf, profiler);
  }
}

// The levels whose reverse matches the forward pass SumProductForward ran
// for these buffers: with grouped compressions only the groups' slots exist.
const std::vector<Level> &ReverseProgram(const Impl &w, const LinearBuffers &buf) {
  return UsesFusedForward(w, buf) && !w.grouped_levels.empty() ? w.grouped_levels : w.levels;
}

// Without branch adjoint storage (buf.branch_adj null) rakes take their
// adjoint from the parent node and the pure-copy levels are skipped.
// `raw_leaves` applies to the first level (round 0's rakes).
void ReverseLevels(const Launch &l, const std::vector<Level> &levels, const LinearBuffers &buf,
                   Profiler *profiler = nullptr, const void *raw_leaves = nullptr) {
  LinearBuffers raw = buf;
  raw.raw_leaves = raw_leaves;
  for (std::size_t i = levels.size(); i-- > 0;) {
    if (buf.branch_adj == nullptr &&
        (levels[i].primitive == THMM_BRANCH_COMBINATION || levels[i].primitive == THMM_BRANCH_ABSORPTION))
      continue;
    const bool grouped = levels[i].primitive == kChainPrimitive || levels[i].primitive == kSmallChainPrimitive ||
                         levels[i].primitive == kPairChainPrimitive;
    if (profiler)
      profiler->Begin(3, levels[i].primitive, grouped ? levels[i].reverse_work : levels[i].count, levels[i].count);
    LaunchReverseLevel(l, levels[i].primitive, levels[i].offset, levels[i].count, i == 0 ? raw : buf);
    if (profiler) profiler->End();
  }
}

// The reverse pass matching SumProductForward's forward pass: one persistent
// launch, or the reverse program's levels.
void RunReverse(const Impl &w, const Launch &l, const LinearBuffers &buf, Profiler *profiler = nullptr,
                const void *raw_leaves = nullptr) {
  if (PersistentReady(w, l, buf)) {
    if (profiler) profiler->Begin(3, kPersistentPrimitive, w.persistent_reverse.num_tasks);
    LaunchPersistentPass(l, true, w.Gather(), QueueOf(w, w.persistent_reverse), buf);
    if (profiler) profiler->End();
    return;
  }
  ReverseLevels(l, ReverseProgram(w, buf), buf, profiler, raw_leaves);
}

// Whether round 0's rakes of a dense call can read the caller's linear node
// potentials `nodes` ([B][N][K], workspace element type) while the node
// initialisation skips their leaves: the workspace has the raw-leaf layout,
// the pass runs the fused program (its first gather step holds round 0's
// rakes) and, for a reverse pass too, the reverse program's first level is
// round 0's rakes on the team kernels (16-byte aligned buffers); the
// initialisation must run on the team transpose.
bool RawLeaves(const Impl &w, const Launch &l, const void *nodes, const LinearBuffers &buf, bool reverse) {
  const auto aligned = [](const void *p) { return (reinterpret_cast<std::uintptr_t>(p) & 15u) == 0; };
  if (w.tuning.input_leaves == 0 || !w.round0_gather || !UsesFusedForward(w, buf) || PersistentReady(w, l, buf) ||
      buf.log_edges != nullptr || buf.edges_f32 != 0)
    return false;
  if (nodes == nullptr || !aligned(nodes) || !aligned(w.nodes.ptr) || !FastTransposeFits(l.dims.K, l.dims.B, l.dims.N))
    return false;
  if (!reverse) return true;
  const std::vector<Level> &program = ReverseProgram(w, buf);
  return !program.empty() && program.front().primitive == THMM_RAKE && program.front().offset == 0 &&
         program.front().count == w.round0_rakes && aligned(buf.edges) && aligned(buf.edge_out) &&
         aligned(buf.node_adj);
}

// Categorical staging with residency policies (src/gpu_backend_impl.inc:2841-2947).
struct CategoricalPolicy {
  bool update_observations;
  bool update_factors;
};

CategoricalPolicy StageCategorical(Impl &w, const thmm_categorical_model &model, thmm_categorical_update update) {
  if (!w.categorical || model.categories != w.categories || model.num_observation_nodes != w.observation_nodes.size() ||
      (model.num_observation_nodes != 0 &&
       (model.observation_nodes == nullptr ||
        !std::equal(model.observation_nodes, model.observation_nodes + model.num_observation_nodes,
                    w.observation_nodes.begin()))))
    throw std::invalid_argument("categorical accelerator model does not match the reserved workspace");
  const std::size_t observation_values = Product({model.batch, model.num_observation_nodes}, "categorical observations");
  const std::size_t emission_values = Product({model.categories, model.states}, "categorical emissions");
  if (model.observation_count != observation_values || model.root_count != model.states ||
      model.emission_count != emission_values)
    throw std::invalid_argument("accelerator categorical input shapes are wrong");
  if (update != THMM_UPDATE_ALL && update != THMM_UPDATE_FACTORS && update != THMM_UPDATE_NONE)
    throw std::invalid_argument("unknown categorical input update policy");
  const CategoricalPolicy policy{update == THMM_UPDATE_ALL, update != THMM_UPDATE_NONE};
  if (!policy.update_observations && w.resident_observation_batch != model.batch)
    throw std::logic_error("categorical observations are not resident for this batch; use kAll");
  if (!policy.update_factors && !w.resident_factors)
    throw std::logic_error("categorical factors are not resident; use kAll or kFactors");
  if (policy.update_observations && model.observations != w.host_observations.ptr && observation_values != 0)
    std::memcpy(w.host_observations.ptr, model.observations, observation_values);
  if (policy.update_factors) {
    if (model.root_potential != w.host_root.ptr) std::memcpy(w.host_root.ptr, model.root_potential, w.states * w.S());
    if (model.emission_potentials != w.host_emissions.ptr)
      std::memcpy(w.host_emissions.ptr, model.emission_potentials, emission_values * w.S());
    if (model.edge_potentials != w.host_edges.ptr && model.edge_count != 0)
      std::memcpy(w.host_edges.ptr, model.edge_potentials, model.edge_count * w.S());
  }
  return policy;
}

void UploadCategorical(Impl &w, const thmm_categorical_model &model, const CategoricalPolicy &policy) {
  if (policy.update_observations && model.observation_count != 0) {
    Hip(hipMemcpyAsync(w.observations.ptr, w.host_observations.ptr, model.observation_count, hipMemcpyHostToDevice,
                       w.stream),
        "GPU runtime MemcpyAsync categorical observation upload");
    LaunchTransposeObservations(w.MakeLaunch(w.real, model.batch, w.stream), w.observations.as<std::uint8_t>(),
                                w.observations_t.as<std::uint8_t>(), static_cast<std::uint32_t>(model.batch),
                                static_cast<std::uint32_t>(model.num_observation_nodes));
  }
  if (policy.update_factors) {
    Hip(hipMemcpyAsync(w.root_potential.ptr, w.host_root.ptr, w.states * w.S(), hipMemcpyHostToDevice, w.stream),
        "GPU runtime MemcpyAsync root-potential upload");
    Hip(hipMemcpyAsync(w.emissions.ptr, w.host_emissions.ptr, model.emission_count * w.S(), hipMemcpyHostToDevice,
                       w.stream),
        "GPU runtime MemcpyAsync emission-potential upload");
    if (model.edge_count != 0)
      Hip(hipMemcpyAsync(w.edges.ptr, w.host_edges.ptr, model.edge_count * w.S(), hipMemcpyHostToDevice, w.stream),
          "GPU runtime MemcpyAsync edge upload");
  }
}

void FinishCategorical(Impl &w, const thmm_categorical_model &model, const CategoricalPolicy &policy) {
  if (policy.update_observations) w.resident_observation_batch = model.batch;
  if (policy.update_factors) w.resident_factors = true;
}

// Where a call's node factors come from: the staged dense potentials or the
// staged categorical observations.
struct Source {
  const thmm_categorical_model *categorical = nullptr;
};

// Whether the categorical node state of launch l takes its logs in the
// element type: exactly when the dense initialisation of the same factors
// (staged, aligned [B][N][K]) does, so that both give the same state.
bool ElementLogs(const Impl &w, const Launch &l) {
  return l.real == Real::kF32 && w.real == Real::kF32 && FastTransposeFits(l.dims.K, l.dims.B, l.dims.N);
}

// Writes the internal node state (max-shifted logs + scales, element type
// l.real) of a call from its staged inputs, or for a dense call from
// `dense_nodes` ([B][N][K], the workspace element type) when given. With
// `fused_leaves` a categorical call leaves the leaves of the workspace's
// categorical rakes unwritten (their rake kernels form them).
void InitState(Impl &w, const Source &src, const Launch &l, void *nodes, double *scale,
               const void *dense_nodes = nullptr, bool fused_leaves = false, const NodeList *dense_only = nullptr) {
  if (src.categorical == nullptr) {
    LaunchInitNodes(l, w.real, dense_nodes != nullptr ? dense_nodes : w.input_nodes.ptr, nodes, scale, dense_only);
  } else {
    LaunchInitCategorical(l, w.real, false, w.observations_t.as<std::uint8_t>(), w.observation_index.as<std::uint32_t>(),
                          static_cast<std::uint32_t>(src.categorical->num_observation_nodes),
                          static_cast<std::uint32_t>(src.categorical->categories), w.is_root.as<std::uint8_t>(),
                          w.root_potential.ptr, w.emissions.ptr, nodes, scale,
                          fused_leaves && w.cat_rakes != 0 && w.tuning.input_leaves != 0
                              ? w.cat_skip.as<std::uint8_t>()
                              : nullptr,
                          ElementLogs(w, l));
  }
}

// The partition / marginal buffers of a call: for a categorical call on a
// workspace with leaf fusion, its first cat_rakes rakes read the staged
// observations, emissions and root potential.
LinearBuffers CallLinear(const Impl &w, const Launch &l, bool categorical) {
  LinearBuffers buf = w.Linear(w.edges.ptr, nullptr);
  if (categorical && w.cat_rakes != 0 && w.tuning.input_leaves != 0)
    buf.cat = {w.observations_t.as<std::uint8_t>(), w.observation_index.as<std::uint32_t>(),
               w.is_root.as<std::uint8_t>(), w.root_potential.ptr, w.emissions.ptr,
               static_cast<std::uint32_t>(w.categories), w.cat_rakes, ElementLogs(w, l) ? 1u : 0u};
  return buf;
}

void DownloadLogZ(Impl &w, std::size_t batch) {
  Hip(hipMemcpyAsync(w.host_log_z.ptr, w.log_z.ptr, batch * w.Roots() * sizeof(double), hipMemcpyDeviceToHost,
                     w.stream),
      "GPU runtime MemcpyAsync output download");
}

void WaitDownload(Impl &w) { Hip(hipEventSynchronize(w.events[5]), "tree-HMM accelerator execution"); }

// Converts the FP64 log partitions to Scalar outputs (exp for Z).
void PublishValues(Impl &w, std::size_t batch, bool log_domain) {
  const double *src = w.host_log_z.as<double>();
  const std::size_t count = batch * w.Roots();
  for (std::size_t i = 0; i < count; ++i) {
    const double v = log_domain ? src[i] : std::exp(src[i]);
    if (w.real == Real::kF32)
      w.host_values.as<float>()[i] = static_cast<float>(v);
    else
      w.host_values.as<double>()[i] = v;
  }
}

void RequireFiniteLogZ(const Impl &w, std::size_t batch, const char *message) {
  const double *src = w.host_log_z.as<double>();
  for (std::size_t i = 0; i < batch * w.Roots(); ++i)
    if (!std::isfinite(src[i])) throw std::domain_error(message);
}

// With one batch item the internal [N][B][K] node layout is the public
// [B][N][K] one, so the reverse pass writes node marginals straight into the
// output instead of transposing them afterwards.
void AdjointsInto(LinearBuffers &buf, std::size_t batch, void *node_out) {
  if (batch == 1) buf.node_adj = node_out;
}

// Scaled forward + log Z for a staged call.
// Round 0's rakes of a dense staged call read the staged linear potentials
// when RawLeaves allows it (the initialisation then skips their leaves).
const void *StagedRawLeaves(const Impl &w, const Launch &l, const Source &src, const LinearBuffers &buf,
                            bool reverse) {
  return src.categorical == nullptr && RawLeaves(w, l, w.input_nodes.ptr, buf, reverse) ? w.input_nodes.ptr : nullptr;
}

void RunPartition(Impl &w, std::size_t batch, const Source &src) {
  const Launch l = w.MakeLaunch(w.real, batch, w.stream);
  const LinearBuffers buf = CallLinear(w, l, src.categorical != nullptr);
  const void *raw = StagedRawLeaves(w, l, src, buf, false);
  const NodeList init = w.InitList();
  InitState(w, src, l, w.nodes.ptr, w.node_scale.as<double>(), nullptr, true,
            raw != nullptr ? &init : nullptr);
  SumProductForward(w, l, buf, nullptr, raw);
  LaunchFinish(l, buf, w.log_z.as<double>());
  CheckLaunches();
}

void RunMarginals(Impl &w, std::size_t batch, const Source &src) {
  const Launch l = w.MakeLaunch(w.real, batch, w.stream);
  LinearBuffers buf = CallLinear(w, l, src.categorical != nullptr);
  AdjointsInto(buf, batch, w.node_out.ptr);
  const void *raw = StagedRawLeaves(w, l, src, buf, true);
  const NodeList init = w.InitList();
  InitState(w, src, l, w.nodes.ptr, w.node_scale.as<double>(), nullptr, true,
            raw != nullptr ? &init : nullptr);
  SumProductForward(w, l, buf, nullptr, raw);
  LaunchFinish(l, buf, w.log_z.as<double>());
  LaunchSeedAdjoint(l, buf, nullptr);
  RunReverse(w, l, buf, nullptr, raw);
  if (batch != 1) LaunchNodeAdjointOut(l, w.node_adj.ptr, w.node_out.ptr);
  CheckLaunches();
}

// Where a MAP or sampling call reads its dense inputs and writes its
// outputs: the staged workspace buffers, or a device-resident caller's.
struct DenseCall {
  const void *nodes, *edges, *uniforms;
  double *log_z;             // log partitions, or MAP log weights
  std::uint32_t *assignment;  // [B][N]
  std::uint32_t *failure;
  hipStream_t stream;
};

DenseCall StagedCall(Impl &w, double *log_z) {
  return {w.input_nodes.ptr, w.edges.ptr, w.uniforms.ptr, log_z, w.assignment.as<std::uint32_t>(),
          w.failure.as<std::uint32_t>(), w.stream};
}

void RunMaximum(Impl &w, std::size_t batch, const Source &src, const DenseCall &call) {
  const Launch l = w.MakeLaunch(w.real, batch, call.stream);
  // At K = 8 and 16 the level kernels log the edge factors on load from
  // 16-byte aligned raw potentials; the log transform then covers the nodes
  // only (a launch without edges).
  const void *raw_edges = src.categorical == nullptr ? call.edges : w.edges.ptr;
  const bool edges_on_load = (w.states == 8 || w.states == 16) && w.E() != 0 &&
                             (reinterpret_cast<std::uintptr_t>(raw_edges) & 15u) == 0;
  Launch init = l;
  if (edges_on_load) init.dims.E = 0;
  const bool pieces = !w.max_levels.empty() && w.tuning.compress_depth >= 2;
  const std::vector<Level> &levels = pieces ? w.max_levels : w.levels;
  // Round 0's leaves are logged on load by the team rake kernels of the
  // first level (dense 16-byte aligned inputs at K = 8, 16 and 32, where that
  // level stays on the team kernels; each lane logs one value of its quad),
  // and round 0's parents by its absorption levels; MaxInit logs the other
  // nodes.
  const bool leaves_on_load =
      src.categorical == nullptr && w.tuning.input_leaves != 0 && w.round0_rakes != 0 &&
      (reinterpret_cast<std::uintptr_t>(call.nodes) & 15u) == 0 &&
      !levels.empty() && levels.front().primitive == THMM_RAKE && levels.front().offset == 0 &&
      levels.front().count == w.round0_rakes &&
      (w.states == 8 || w.states == 16 ||
       (w.states == 32 && static_cast<std::uint64_t>(w.round0_rakes) * batch * 32 * 32 / 4 < kMaxLaunchItems));
  const NodeList inner = w.InitList();
  // Logs of the raw staged factors (bit-exact with the oracle's log of the
  // same values), internal layout.
  if (src.categorical == nullptr) {
    LaunchMaxInit(init, w.real, call.nodes, call.edges, w.m_nodes.as<double>(), w.m_edges.as<double>(),
                  leaves_on_load ? &inner : nullptr);
  } else {
    LaunchInitCategorical(l, w.real, true, w.observations_t.as<std::uint8_t>(), w.observation_index.as<std::uint32_t>(),
                          static_cast<std::uint32_t>(src.categorical->num_observation_nodes),
                          static_cast<std::uint32_t>(src.categorical->categories), w.is_root.as<std::uint8_t>(),
                          w.root_potential.ptr, w.emissions.ptr, w.nodes.ptr, nullptr);
    // The raw products are internal [N][B][K]: present them to MaxInit as a
    // [1][N*B][K] "public" array so that its transpose is the identity.
    Launch as_public = init;
    as_public.dims.B = 1;
    as_public.dims.N = static_cast<std::uint32_t>(w.N() * batch);
    LaunchMaxInit(as_public, w.real, w.nodes.ptr, w.edges.ptr, w.m_nodes.as<double>(), w.m_edges.as<double>());
  }
  MaxBuffers buf = w.Max();
  if (edges_on_load) {
    buf.raw_edges = raw_edges;
    buf.raw_edges_f32 = w.real == Real::kF32 ? 1 : 0;
  }
  // The pieces program's direct absorptions: its rake kernels absorb the
  // branches no combination touches, and its absorption levels index the
  // others in max_absorbs.
  Launch program = l;
  MaxBuffers forward = buf;
  if (pieces && w.max_direct.ptr != nullptr) {
    forward.rake_direct = w.max_direct.as<std::uint8_t>();
    program.ops.absorbs = w.max_absorbs.as<DeviceAbsorb>();
  }
  MaxBuffers raw = forward;
  raw.raw_leaves = call.nodes;
  raw.raw_leaves_f32 = w.real == Real::kF32 ? 1 : 0;
  bool round0 = true;
  for (std::size_t i = 0; i < levels.size(); ++i) {
    const Level &level = levels[i];
    round0 = round0 && (i == 0 || level.primitive >= levels[i - 1].primitive);
    const bool from_raw = leaves_on_load && round0 &&
                          (level.primitive == THMM_RAKE ? i == 0 : level.primitive == THMM_BRANCH_ABSORPTION);
    LaunchMaxLevel(program, level.primitive, level.offset, level.count, from_raw ? raw : forward);
  }
  LaunchMaxFinish(l, buf, call.log_z, call.assignment);
  for (std::size_t i = w.levels.size(); i-- > 0;)
    LaunchMaxExpandLevel(l, w.levels[i].primitive, w.levels[i].offset, w.levels[i].count, buf, call.assignment);
  CheckLaunches();
}

void RunMaximum(Impl &w, std::size_t batch, const Source &src) {
  RunMaximum(w, batch, src, StagedCall(w, w.log_weight.as<double>()));
}

void RunSample(Impl &w, std::size_t batch, const Source &src, const DenseCall &call) {
  const Launch l = w.MakeLaunch(Real::kF64, batch, call.stream);
  // At K = 8 and 16 the FP64 level kernels widen 16-byte aligned FP32 edges
  // as they load them; otherwise the edges are widened into s_edges first.
  const bool widen_on_load = w.real == Real::kF32 && (w.states == 8 || w.states == 16) &&
                             (reinterpret_cast<std::uintptr_t>(call.edges) & 15u) == 0;
  if (w.real == Real::kF32) {
    // FP64 sampling state from the float factors.
    InitState(w, src, l, w.s_nodes.ptr, w.s_node_scale.as<double>(), call.nodes);
    if (!widen_on_load)
      LaunchWidenF32(l, static_cast<const float *>(call.edges), w.s_edges.as<double>(), w.E() * w.KK());
  } else {
    InitState(w, src, l, w.nodes.ptr, w.node_scale.as<double>(), call.nodes);
  }
  LaunchClearWord(l, call.failure);
  LinearBuffers buf = w.SamplingLinear(call.edges);
  if (widen_on_load) {
    buf.edges = call.edges;
    buf.edges_f32 = 1;
  }
  // The fused program with plain compression levels (its gather kernels
  // widen FP32 edges on load too) up to K = 32. At K = 64 the FP64 gather
  // holds 216 registers (two wavefronts per SIMD) and measured no faster
  // than the plan's levels.
  if (w.states <= 32 && UsesFusedForward(w, buf))
    RunFusedSteps(w, w.fused_plain, l, buf);
  else
    ForwardLevels(l, w.levels, buf);
  LaunchFinish(l, buf, call.log_z);
  LaunchSampleSeed(l, buf, w.real, call.uniforms, call.assignment, call.failure);
  for (std::size_t i = w.levels.size(); i-- > 0;)
    LaunchSampleExpandLevel(l, w.levels[i].primitive, w.levels[i].offset, w.levels[i].count, buf, w.real,
                            call.uniforms, call.assignment, call.failure);
  CheckLaunches();
}

void RunSample(Impl &w, std::size_t batch, const Source &src) {
  RunSample(w, batch, src, StagedCall(w, w.log_z.as<double>()));
}

void ValidateUniforms(const Impl &w, const void *uniforms, std::size_t count, std::size_t batch) {
  if (!(w.flags & THMM_RESERVE_SAMPLING) || count != Product({batch, w.N()}, "posterior assignments"))
    throw std::invalid_argument(
        "prepared accelerator posterior sampling requires ReserveSampling and one uniform variate per batch item "
        "and node");
  if (count != 0 && uniforms == nullptr) throw std::invalid_argument("uniforms are null");
  // src/gpu_backend_impl.inc:2182-2187
  for (std::size_t i = 0; i < count; ++i) {
    const double u = w.real == Real::kF32 ? static_cast<const float *>(uniforms)[i]
                                          : static_cast<const double *>(uniforms)[i];
    if (!std::isfinite(u) || u < 0.0 || u >= 1.0)
      throw std::invalid_argument("posterior-sampling variates must lie in [0, 1)");
  }
}

}  // namespace

Workspace::Workspace() : impl_(std::make_unique<Impl>()) {}
Workspace::~Workspace() = default;

void Workspace::Reserve(const Plan &plan, const void *identity, std::size_t states, std::size_t batch, Real real,
                        unsigned flags, int device) {
  Impl &w = *impl_;
  try {
    ReserveCommon(w, plan, identity, states, batch, real, flags, device);
    const std::size_t node_values = Product({batch, plan.num_nodes, states}, "node inputs");
    if (!w.device_resident()) {
      w.input_nodes.Allocate(node_values * w.S(), "node inputs");
      w.host_nodes.Allocate(node_values * w.S(), "pinned node inputs");
    }
  } catch (...) {
    w.Clear();
    throw;
  }
}

void Workspace::ReserveCategorical(const Plan &plan, const void *identity, std::size_t states, std::size_t batch,
                                   std::size_t categories, const std::uint32_t *observation_nodes,
                                   std::size_t observation_count, Real real, unsigned flags, int device) {
  // src/gpu_backend_impl.inc:1985-1998
  if (categories == 0 || categories > 256)
    throw std::invalid_argument("accelerator categorical model must have between 1 and 256 categories");
  if (observation_count != 0 && observation_nodes == nullptr) throw std::invalid_argument("observation nodes are null");
  for (std::size_t i = 0; i < observation_count; ++i)
    if (observation_nodes[i] >= plan.num_nodes || (i > 0 && observation_nodes[i] <= observation_nodes[i - 1]))
      throw std::invalid_argument(
          "accelerator categorical observation nodes must be valid and strictly increasing");
  Impl &w = *impl_;
  try {
    ReserveCommon(w, plan, identity, states, batch, real, flags & ~kReserveDeviceResident, device);
    w.categorical = true;
    w.categories = categories;
    w.observation_nodes.assign(observation_nodes, observation_nodes + observation_count);
    const std::size_t observation_values = Product({batch, observation_count}, "categorical observations");
    const std::size_t emission_values = Product({categories, states}, "categorical emissions");
    w.host_observations.Allocate(observation_values, "pinned observations");
    w.host_root.Allocate(states * w.S(), "pinned root potential");
    w.host_emissions.Allocate(emission_values * w.S(), "pinned emissions");
    w.observations.Allocate(observation_values, "observations");
    w.observations_t.Allocate(observation_values, "observations");
    w.root_potential.Allocate(states * w.S(), "root potential");
    w.emissions.Allocate(emission_values * w.S(), "emissions");
    if (flags & THMM_RESERVE_MARGINALS) {
      const std::size_t seeds = batch * plan.roots.size(), E = plan.num_edges;
      w.chain_seed.Allocate(seeds * sizeof(double), "branch-length gradient seeds");
      w.chain_rate.Allocate(states * states * sizeof(double), "rate matrix");
      w.chain_terms.Allocate(E * states * states * sizeof(double), "branch-length chain terms");
      w.chain_grad.Allocate(E * sizeof(double), "branch-length gradients");
      w.host_chain_seed.Allocate(seeds * sizeof(double), "pinned gradient seeds");
      w.host_chain_rate.Allocate(states * states * sizeof(double), "pinned rate matrix");
      w.host_chain_grad.Allocate(E * sizeof(double), "pinned branch-length gradients");
    }
    std::vector<std::uint32_t> index(plan.num_nodes, 0xFFFFFFFFu);
    for (std::size_t i = 0; i < observation_count; ++i) index[observation_nodes[i]] = static_cast<std::uint32_t>(i);
    Upload(w.observation_index, index, w.stream, "categorical topology upload");
    // Leaf fusion where the levels run on the generic rake kernels (no fused
    // forward program, K other than 8, 16, 32, 64): the longest prefix of the
    // rakes whose leaves have no children. Each such leaf is read by its
    // rake alone, forward and reverse.
    if (w.fused.empty() && states != 8 && states != 16 && states != 32 && states != 64) {
      std::vector<std::uint8_t> has_child(plan.num_nodes, 0), skip(plan.num_nodes, 0);
      for (Index parent : plan.edge_parents) has_child[parent] = 1;
      std::size_t r = 0;
      while (r < plan.rakes.size() && !has_child[plan.rakes[r].leaf]) skip[plan.rakes[r++].leaf] = 1;
      if (r != 0) {
        Upload(w.cat_skip, skip, w.stream, "categorical topology upload");
        w.cat_rakes = static_cast<std::uint32_t>(r);
      }
    }
    w.site_sums.Allocate(plan.roots.size() * sizeof(double), "log-likelihood sums");
    w.host_site_sums.Allocate(plan.roots.size() * sizeof(double), "pinned log-likelihood sums");
    Hip(hipStreamSynchronize(w.stream), "categorical topology upload");
  } catch (...) {
    w.Clear();
    throw;
  }
}

void Workspace::Inputs(std::size_t batch, void **nodes, std::size_t *node_count, void **edges,
                       std::size_t *edge_count) {
  Impl &w = *impl_;
  if (w.plan == nullptr) throw std::logic_error("accelerator Workspace::Reserve must precede Inputs");
  if (w.categorical) throw std::logic_error("accelerator dense Inputs cannot be used after ReserveCategorical");
  if (w.device_resident()) throw std::logic_error("this workspace has no pinned input storage");
  if (batch == 0 || batch > w.capacity) throw std::invalid_argument("accelerator input batch exceeds the reserved capacity");
  *nodes = w.host_nodes.ptr;
  *node_count = Product({batch, w.N(), w.states}, "accelerator input nodes");
  *edges = w.host_edges.ptr;
  *edge_count = w.E() * w.KK();
}

void Workspace::CategoricalInputs(std::size_t batch, std::uint8_t **observations, std::size_t *observation_count,
                                  void **root, void **emissions, std::size_t *emission_count, void **edges,
                                  std::size_t *edge_count) {
  Impl &w = *impl_;
  if (w.plan == nullptr || !w.categorical)
    throw std::logic_error("accelerator Workspace::ReserveCategorical must precede CategoricalInputs");
  if (batch == 0 || batch > w.capacity)
    throw std::invalid_argument("accelerator categorical input batch exceeds the reserved capacity");
  *observations = w.host_observations.as<std::uint8_t>();
  *observation_count = Product({batch, w.observation_nodes.size()}, "categorical observations");
  *root = w.host_root.ptr;
  *emissions = w.host_emissions.ptr;
  *emission_count = w.categories * w.states;
  *edges = w.host_edges.ptr;
  *edge_count = w.E() * w.KK();
}

void Workspace::Uniforms(std::size_t batch, void **uniforms, std::size_t *count) {
  Impl &w = *impl_;
  if (!(w.flags & THMM_RESERVE_SAMPLING) || w.plan == nullptr)
    throw std::logic_error("accelerator Workspace::ReserveSampling must precede Uniforms");
  if (w.device_resident()) throw std::logic_error("this workspace has no pinned uniform storage");
  if (batch == 0 || batch > w.capacity) throw std::invalid_argument("accelerator uniform batch exceeds the reserved capacity");
  *uniforms = w.host_uniforms.ptr;
  *count = Product({batch, w.N()}, "accelerator posterior uniforms");
}

void Workspace::Partition(const thmm_dense_model &model, bool log_domain, const void **values, std::size_t *count,
                          thmm_timings *timings) {
  Impl &w = *impl_;
  RequireHost(w);
  const Timer timer;
  CheckCall(w, model.identity, model.states, model.batch, model.edge_count, "Workspace::Reserve");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  StageDense(w, model);
  Record(w, 0);
  UploadDense(w, model);
  Record(w, 1);
  Record(w, 2);
  RunPartition(w, model.batch, Source{});
  Record(w, 3);
  Record(w, 4);
  DownloadLogZ(w, model.batch);
  Record(w, 5);
  WaitDownload(w);
  PublishValues(w, model.batch, log_domain);
  *values = w.host_values.ptr;
  *count = model.batch * w.Roots();
  FillTimings(w, timer, timings);
}

void Workspace::PartitionCategorical(const thmm_categorical_model &model, thmm_categorical_update update,
                                     bool log_domain, const void **values, std::size_t *count,
                                     thmm_timings *timings) {
  Impl &w = *impl_;
  RequireHost(w);
  const Timer timer;
  CheckCall(w, model.identity, model.states, model.batch, model.edge_count, "Workspace::Reserve");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  const CategoricalPolicy policy = StageCategorical(w, model, update);
  Record(w, 0);
  UploadCategorical(w, model, policy);
  Record(w, 1);
  Record(w, 2);
  RunPartition(w, model.batch, Source{&model});
  Record(w, 3);
  Record(w, 4);
  DownloadLogZ(w, model.batch);
  Record(w, 5);
  WaitDownload(w);
  FinishCategorical(w, model, policy);
  PublishValues(w, model.batch, log_domain);
  *values = w.host_values.ptr;
  *count = model.batch * w.Roots();
  FillTimings(w, timer, timings);
}

namespace {

void DownloadMaximum(Impl &w, std::size_t batch, const void **log_weights, const std::uint32_t **states) {
  Hip(hipMemcpyAsync(w.host_log_z.ptr, w.log_weight.ptr, batch * w.Roots() * sizeof(double), hipMemcpyDeviceToHost,
                     w.stream),
      "GPU runtime MemcpyAsync MAP-weight download");
  Hip(hipMemcpyAsync(w.host_assignment.ptr, w.assignment.ptr, batch * w.N() * sizeof(std::uint32_t),
                     hipMemcpyDeviceToHost, w.stream),
      "GPU runtime MemcpyAsync MAP-assignment download");
  Record(w, 5);
  WaitDownload(w);
  // src/gpu_backend_impl.inc:2581-2585
  RequireFiniteLogZ(w, batch, "the tree HMM has no positive-weight assignment");
  PublishValues(w, batch, true);
  *log_weights = w.host_values.ptr;
  *states = w.host_assignment.as<std::uint32_t>();
}

void RequireMaximum(const Impl &w) {
  if (!(w.flags & THMM_RESERVE_MAXIMUM))
    throw std::invalid_argument(
        "prepared accelerator MAP inference requires ReserveMaximum for this plan, state count, and batch "
        "capacity");
}

void RequireMarginals(const Impl &w) {
  if (!(w.flags & THMM_RESERVE_MARGINALS))
    throw std::invalid_argument(
        "prepared accelerator marginal inference requires ReserveMarginals for this plan, state count, and batch "
        "capacity");
}

void DownloadSample(Impl &w, std::size_t batch, const std::uint32_t **states) {
  Hip(hipMemcpyAsync(w.host_assignment.ptr, w.assignment.ptr, batch * w.N() * sizeof(std::uint32_t),
                     hipMemcpyDeviceToHost, w.stream),
      "GPU runtime MemcpyAsync posterior-sample download");
  Hip(hipMemcpyAsync(w.host_failure.ptr, w.failure.ptr, sizeof(std::uint32_t), hipMemcpyDeviceToHost, w.stream),
      "GPU runtime MemcpyAsync sampling status download");
  Record(w, 5);
  WaitDownload(w);
  // src/gpu_backend_impl.inc:2377-2384
  RequireFiniteLogZ(w, batch, "the tree HMM has a nonpositive partition function");
  if (*w.host_failure.as<std::uint32_t>() != 0)
    throw std::domain_error("cannot sample from an empty conditional distribution");
  *states = w.host_assignment.as<std::uint32_t>();
}

void DownloadMarginals(Impl &w, std::size_t batch, const void **log_partitions, const void **nodes,
                       const void **edges) {
  const std::size_t node_values = batch * w.N() * w.states;
  const std::size_t edge_values = batch * w.E() * w.KK();
  Hip(hipMemcpyAsync(w.host_node_out.ptr, w.node_out.ptr, node_values * w.S(), hipMemcpyDeviceToHost, w.stream),
      "GPU runtime MemcpyAsync node-marginal download");
  if (edge_values != 0)
    Hip(hipMemcpyAsync(w.host_edge_out.ptr, w.edge_out.ptr, edge_values * w.S(), hipMemcpyDeviceToHost, w.stream),
        "GPU runtime MemcpyAsync edge-marginal download");
  Record(w, 5);
  WaitDownload(w);
  // src/gpu_backend_impl.inc:2780-2785
  RequireFiniteLogZ(w, batch, "the tree HMM has a nonpositive partition function");
  PublishValues(w, batch, true);
  *log_partitions = w.host_values.ptr;
  *nodes = w.host_node_out.ptr;
  *edges = w.host_edge_out.ptr;
}

void UploadUniforms(Impl &w, const void *uniforms, std::size_t count) {
  if (uniforms != w.host_uniforms.ptr) std::memcpy(w.host_uniforms.ptr, uniforms, count * w.S());
  Hip(hipMemcpyAsync(w.uniforms.ptr, w.host_uniforms.ptr, count * w.S(), hipMemcpyHostToDevice, w.stream),
      "GPU runtime MemcpyAsync posterior-uniform upload");
}

}  // namespace

void Workspace::Maximum(const thmm_dense_model &model, const void **log_weights, const std::uint32_t **states,
                        thmm_timings *timings) {
  Impl &w = *impl_;
  RequireHost(w);
  const Timer timer;
  RequireMaximum(w);
  CheckCall(w, model.identity, model.states, model.batch, model.edge_count, "ReserveMaximum");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  StageDense(w, model);
  Record(w, 0);
  UploadDense(w, model);
  Record(w, 1);
  Record(w, 2);
  RunMaximum(w, model.batch, Source{});
  Record(w, 3);
  Record(w, 4);
  DownloadMaximum(w, model.batch, log_weights, states);
  FillTimings(w, timer, timings);
}

void Workspace::MaximumCategorical(const thmm_categorical_model &model, thmm_categorical_update update,
                                   const void **log_weights, const std::uint32_t **states, thmm_timings *timings) {
  Impl &w = *impl_;
  RequireHost(w);
  const Timer timer;
  RequireMaximum(w);
  CheckCall(w, model.identity, model.states, model.batch, model.edge_count, "ReserveCategoricalMaximum");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  const CategoricalPolicy policy = StageCategorical(w, model, update);
  Record(w, 0);
  UploadCategorical(w, model, policy);
  Record(w, 1);
  Record(w, 2);
  RunMaximum(w, model.batch, Source{&model});
  Record(w, 3);
  Record(w, 4);
  DownloadMaximum(w, model.batch, log_weights, states);
  FinishCategorical(w, model, policy);
  FillTimings(w, timer, timings);
}

void Workspace::Sample(const thmm_dense_model &model, const void *uniforms, std::size_t uniform_count,
                       const std::uint32_t **states, thmm_timings *timings) {
  Impl &w = *impl_;
  RequireHost(w);
  const Timer timer;
  CheckCall(w, model.identity, model.states, model.batch, model.edge_count, "ReserveSampling");
  ValidateUniforms(w, uniforms, uniform_count, model.batch);
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  StageDense(w, model);
  Record(w, 0);
  UploadDense(w, model);
  UploadUniforms(w, uniforms, uniform_count);
  Record(w, 1);
  Record(w, 2);
  RunSample(w, model.batch, Source{});
  Record(w, 3);
  Record(w, 4);
  DownloadLogZ(w, model.batch);
  DownloadSample(w, model.batch, states);
  FillTimings(w, timer, timings);
}

void Workspace::SampleCategorical(const thmm_categorical_model &model, const void *uniforms,
                                  std::size_t uniform_count, thmm_categorical_update update,
                                  const std::uint32_t **states, thmm_timings *timings) {
  Impl &w = *impl_;
  RequireHost(w);
  const Timer timer;
  CheckCall(w, model.identity, model.states, model.batch, model.edge_count, "ReserveCategoricalSampling");
  ValidateUniforms(w, uniforms, uniform_count, model.batch);
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  const CategoricalPolicy policy = StageCategorical(w, model, update);
  Record(w, 0);
  UploadCategorical(w, model, policy);
  UploadUniforms(w, uniforms, uniform_count);
  Record(w, 1);
  Record(w, 2);
  RunSample(w, model.batch, Source{&model});
  Record(w, 3);
  Record(w, 4);
  DownloadLogZ(w, model.batch);
  DownloadSample(w, model.batch, states);
  FinishCategorical(w, model, policy);
  FillTimings(w, timer, timings);
}

void Workspace::Marginals(const thmm_dense_model &model, const void **log_partitions, const void **nodes,
                          const void **edges, thmm_timings *timings) {
  Impl &w = *impl_;
  RequireHost(w);
  const Timer timer;
  RequireMarginals(w);
  CheckCall(w, model.identity, model.states, model.batch, model.edge_count, "ReserveMarginals");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  StageDense(w, model);
  Record(w, 0);
  UploadDense(w, model);
  Record(w, 1);
  Record(w, 2);
  RunMarginals(w, model.batch, Source{});
  Record(w, 3);
  Record(w, 4);
  DownloadLogZ(w, model.batch);
  DownloadMarginals(w, model.batch, log_partitions, nodes, edges);
  FillTimings(w, timer, timings);
}

void Workspace::MarginalsCategorical(const thmm_categorical_model &model, thmm_categorical_update update,
                                     const void **log_partitions, const void **nodes, const void **edges,
                                     thmm_timings *timings) {
  Impl &w = *impl_;
  RequireHost(w);
  const Timer timer;
  RequireMarginals(w);
  CheckCall(w, model.identity, model.states, model.batch, model.edge_count, "ReserveCategoricalMarginals");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  const CategoricalPolicy policy = StageCategorical(w, model, update);
  Record(w, 0);
  UploadCategorical(w, model, policy);
  Record(w, 1);
  Record(w, 2);
  RunMarginals(w, model.batch, Source{&model});
  Record(w, 3);
  Record(w, 4);
  DownloadLogZ(w, model.batch);
  DownloadMarginals(w, model.batch, log_partitions, nodes, edges);
  FinishCategorical(w, model, policy);
  FillTimings(w, timer, timings);
}

namespace {

// The forward half of the branch-length pair; with `summed` the log
// partitions are summed over the batch on the device and only the [roots]
// sums are downloaded (into host_site_sums).
void CategoricalForward(Impl &w, const thmm_categorical_model &model, thmm_categorical_update update, bool summed) {
  RequireHost(w);
  RequireMarginals(w);
  if (!w.categorical) throw std::logic_error("the branch-length pair needs a ReserveCategoricalMarginals workspace");
  CheckCall(w, model.identity, model.states, model.batch, model.edge_count, "Workspace::ReserveCategoricalMarginals");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  const CategoricalPolicy policy = StageCategorical(w, model, update);
  UploadCategorical(w, model, policy);
  RunPartition(w, model.batch, Source{&model});
  if (summed) {
    const Launch l = w.MakeLaunch(w.real, model.batch, w.stream);
    LaunchSumOverBatch(l, w.log_z.as<double>(), w.site_sums.as<double>());
    CheckLaunches();
    Hip(hipMemcpyAsync(w.host_site_sums.ptr, w.site_sums.ptr, w.Roots() * sizeof(double), hipMemcpyDeviceToHost,
                       w.stream),
        "GPU runtime MemcpyAsync output download");
  } else {
    DownloadLogZ(w, model.batch);
  }
  Record(w, 5);
  WaitDownload(w);
  FinishCategorical(w, model, policy);
  if (summed) {
    // A sum is finite exactly when every term is.
    for (std::size_t r = 0; r < w.Roots(); ++r)
      if (!std::isfinite(w.host_site_sums.as<double>()[r]))
        throw std::domain_error("the tree HMM has a nonpositive partition function");
  } else {
    RequireFiniteLogZ(w, model.batch, "the tree HMM has a nonpositive partition function");
  }
  w.last_batch = model.batch;
  w.forward = Impl::Forward::kCategorical;
}

}  // namespace

void Workspace::CategoricalLogZForward(const thmm_categorical_model &model, thmm_categorical_update update,
                                       const double **log_z, std::size_t *count) {
  Impl &w = *impl_;
  CategoricalForward(w, model, update, false);
  *log_z = w.host_log_z.as<double>();
  *count = model.batch * w.Roots();
}

void Workspace::CategoricalLogLikelihood(const thmm_categorical_model &model, thmm_categorical_update update,
                                         const double **sums, std::size_t *count) {
  Impl &w = *impl_;
  CategoricalForward(w, model, update, true);
  *sums = w.host_site_sums.as<double>();
  *count = w.Roots();
}

void Workspace::SetBranchLengths(const double *rate_matrix, const double *stationary, const double *branch_lengths,
                                 std::size_t edge_count) {
  Impl &w = *impl_;
  if (w.plan == nullptr) throw std::logic_error("accelerator Workspace::Reserve must precede this call");
  if (!w.categorical) throw std::logic_error("branch lengths set the edge factors of a categorical workspace");
  if (edge_count != w.E()) throw std::invalid_argument("one branch length per edge is required");
  const ReversibleRates rates = DecomposeReversible(rate_matrix, stationary, w.states);
  CheckBranchLengths(branch_lengths, edge_count);
  if (!w.resident_factors)
    throw std::logic_error("categorical factors are not resident; stage them with kAll or kFactors first");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  const std::size_t K = w.states, factors = 2 * K * K + K;
  if (w.expm_factors.bytes < (factors + w.E()) * sizeof(double)) {
    w.expm_factors.Allocate((factors + w.E()) * sizeof(double), "expm factors");
    w.host_expm_factors.Allocate((factors + w.E()) * sizeof(double), "pinned expm factors");
  }
  double *host = w.host_expm_factors.as<double>();
  std::copy(rates.V.begin(), rates.V.end(), host);
  std::copy(rates.lambda.begin(), rates.lambda.end(), host + K * K);
  std::copy(rates.V_inv.begin(), rates.V_inv.end(), host + K * K + K);
  std::copy(branch_lengths, branch_lengths + edge_count, host + factors);
  Hip(hipMemcpyAsync(w.expm_factors.ptr, host, (factors + edge_count) * sizeof(double), hipMemcpyHostToDevice,
                     w.stream),
      "GPU runtime MemcpyAsync expm factor upload");
  const Launch l = w.MakeLaunch(w.real, w.capacity, w.stream);
  const double *device = w.expm_factors.as<double>();
  LaunchExpmEdges(l, device, device + K * K, device + K * K + K, device + factors, w.edges.ptr);
  CheckLaunches();
  Hip(hipStreamSynchronize(w.stream), "tree-HMM accelerator execution");
}

void Workspace::BranchLengthBackward(const double *grad_log_z, const double *rate_matrix, const double **grad,
                                     std::size_t *count) {
  Impl &w = *impl_;
  if (w.forward != Impl::Forward::kCategorical)
    throw std::logic_error("thmm_branch_length_backward requires a preceding thmm_categorical_logz_forward call");
  if (rate_matrix == nullptr) throw std::invalid_argument("rate matrix is null");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  const std::size_t seeds = w.last_batch * w.Roots();
  if (grad_log_z != nullptr) {
    std::memcpy(w.host_chain_seed.ptr, grad_log_z, seeds * sizeof(double));
    Hip(hipMemcpyAsync(w.chain_seed.ptr, w.host_chain_seed.ptr, seeds * sizeof(double), hipMemcpyHostToDevice,
                       w.stream),
        "GPU runtime MemcpyAsync gradient-seed upload");
  }
  std::memcpy(w.host_chain_rate.ptr, rate_matrix, w.KK() * sizeof(double));
  Hip(hipMemcpyAsync(w.chain_rate.ptr, w.host_chain_rate.ptr, w.KK() * sizeof(double), hipMemcpyHostToDevice,
                     w.stream),
      "GPU runtime MemcpyAsync rate-matrix upload");
  const Launch l = w.MakeLaunch(w.real, w.last_batch, w.stream);
  LinearBuffers buf = CallLinear(w, l, true);
  AdjointsInto(buf, w.last_batch, w.node_out.ptr);
  LaunchSeedAdjoint(l, buf, grad_log_z != nullptr ? w.chain_seed.as<double>() : nullptr);
  RunReverse(w, l, buf);
  LaunchBranchLengthChain(l, w.edge_out.ptr, w.edges.ptr, w.chain_rate.as<double>(), w.chain_terms.as<double>(),
                          w.chain_grad.as<double>());
  CheckLaunches();
  Hip(hipMemcpyAsync(w.host_chain_grad.ptr, w.chain_grad.ptr, w.E() * sizeof(double), hipMemcpyDeviceToHost,
                     w.stream),
      "GPU runtime MemcpyAsync branch-length gradient download");
  Record(w, 5);
  WaitDownload(w);
  *grad = w.host_chain_grad.as<double>();
  *count = w.E();
}

void Workspace::LogPartitionsF64(const double **values, std::size_t *count) {
  Impl &w = *impl_;
  RequireHost(w);
  *values = w.host_log_z.as<double>();
  *count = w.capacity * w.Roots();
}

namespace {

hipStream_t ChooseStream(const Impl &w, void *stream) {
  if (stream == THMM_HIP_NULL_STREAM) return nullptr;  // HIP's null stream
  return stream != nullptr ? static_cast<hipStream_t>(stream) : w.stream;
}

// Runs `enqueue` on stream s, or, with THMM_RESERVE_GRAPH, replays a hipGraph
// of it on s, capturing it (on the workspace stream) the first time and
// whenever the key changes. The captured launches contain no memory set or
// copy nodes.
template <typename F>
void RunOrReplay(Impl &w, int kind, const Impl::GraphKey &key, hipStream_t s, const F &enqueue) {
  if (!(w.flags & kReserveGraph)) {
    enqueue(s);
    CheckLaunches();
    return;
  }
  Impl::GraphCache &cache = w.graphs[kind];
  if (cache.exec == nullptr || !(cache.key == key)) {
    cache.Release();
    Hip(hipStreamBeginCapture(w.stream, hipStreamCaptureModeThreadLocal), "GPU runtime StreamBeginCapture");
    enqueue(w.stream);
    hipGraph_t graph = nullptr;
    const hipError_t ended = hipStreamEndCapture(w.stream, &graph);
    CheckLaunches();
    Hip(ended, "GPU runtime StreamEndCapture");
    const hipError_t made = hipGraphInstantiate(&cache.exec, graph, nullptr, nullptr, 0);
    static_cast<void>(hipGraphDestroy(graph));
    if (made != hipSuccess) cache.exec = nullptr;
    Hip(made, "GPU runtime GraphInstantiate");
    cache.key = key;
  }
  Hip(hipGraphLaunch(cache.exec, s), "GPU runtime GraphLaunch");
}

// The launch sequence of thmm_marginals_device on stream `on`.
void EnqueueMarginals(Impl &w, hipStream_t on, const void *nodes, const void *edges, std::size_t batch,
                      double *log_z, void *node_out, void *edge_out, Profiler *p) {
  const Launch l = w.MakeLaunch(w.real, batch, on);
  LinearBuffers buf = w.Linear(edges, nullptr);
  buf.edge_out = edge_out;
  AdjointsInto(buf, batch, node_out);
  const void *raw = RawLeaves(w, l, nodes, buf, true) ? nodes : nullptr;
  if (p) p->Begin(0, 0, 0);
  const NodeList init = w.InitList();
  LaunchInitNodes(l, w.real, nodes, w.nodes.ptr, w.node_scale.as<double>(), raw != nullptr ? &init : nullptr);
  if (p) p->End();
  SumProductForward(w, l, buf, p, raw);
  if (p) p->Begin(2, 0, 0);
  LaunchFinish(l, buf, log_z);
  LaunchSeedAdjoint(l, buf, nullptr);
  if (p) p->End();
  RunReverse(w, l, buf, p, raw);
  if (p) p->Begin(4, 0, 0);
  if (batch != 1) LaunchNodeAdjointOut(l, w.node_adj.ptr, node_out);
  if (p) p->End();
}

// Byte offset of a group's first value in a forest-wide array of per-node
// (K values), per-edge (K * K) or per-root (one) records.
const char *At(const void *base, std::size_t first, std::size_t values, std::size_t size) {
  return static_cast<const char *>(base) + first * values * size;
}
char *At(void *base, std::size_t first, std::size_t values, std::size_t size) {
  return static_cast<char *>(base) + first * values * size;
}

// thmm_marginals_device's launches, group after group for a workspace in
// tree groups (each group's forward, log Z, seed and reverse in turn, so
// nothing is recomputed).
void EnqueueMarginalsAll(Impl &w, hipStream_t on, const void *nodes, const void *edges, std::size_t batch,
                         double *log_z, void *node_out, void *edge_out, Profiler *p) {
  if (w.groups.empty()) {
    EnqueueMarginals(w, on, nodes, edges, batch, log_z, node_out, edge_out, p);
    return;
  }
  const std::size_t K = w.states, S = w.S(), out_s = w.half_edges() ? 2 : S;
  for (Impl::TreeGroup &g : w.groups)
    EnqueueMarginals(*g.w, on, At(nodes, g.first_node, K, S), At(edges, g.first_edge, K * K, S), 1,
                     log_z + g.first_root, At(node_out, g.first_node, K, S), At(edge_out, g.first_edge, K * K, out_s),
                     p);
}

// A persistent pass that stopped waiting on a dependency (its spin limit)
// leaves its failure word set: reported once, then cleared. Blocks until
// the device is idle.
void CheckPersistent(Impl &w) {
  if (w.persistent_failure.ptr == nullptr) return;
  Hip(hipDeviceSynchronize(), "tree-HMM accelerator execution");
  std::uint32_t failed = 0;
  Hip(hipMemcpy(&failed, w.persistent_failure.ptr, sizeof(failed), hipMemcpyDeviceToHost), "persistent pass status");
  if (failed == 0) return;
  Hip(hipMemset(w.persistent_failure.ptr, 0, sizeof(failed)), "persistent pass status");
  throw std::runtime_error("persistent pass: a dependency wait exceeded its spin limit; its results are invalid");
}

void ValidateTuning(const Tuning &t) {
  if (t.compress_items_forward < 1 || t.compress_items_forward > 2 || t.compress_items_reverse < 1 ||
      t.compress_items_reverse > 2)
    throw std::invalid_argument("compressions per wavefront must be 1 or 2");
  if (t.rake_block != 64 && t.rake_block != 128 && t.rake_block != 256)
    throw std::invalid_argument("rake block size must be 64, 128 or 256");
  if (t.gather_chunk < 1 || t.gather_chunk > 64) throw std::invalid_argument("gather chunk must be in [1, 64]");
  if (t.compress_depth < 1 || t.compress_depth > 3) throw std::invalid_argument("compression depth must be 1, 2 or 3");
  if (t.rake_log_edges > 1) throw std::invalid_argument("rake_log_edges must be 0 or 1");
  if (t.persistent > 1) throw std::invalid_argument("persistent must be 0 or 1");
  if (t.input_leaves > 1) throw std::invalid_argument("input_leaves must be 0 or 1");
  if (t.persistent != 0 && !kPersistentPassesEnabled)
    throw std::logic_error(
        "persistent passes are switched off: on the MI355X their launches did not complete (per-task device-wide "
        "release and acquire across the XCDs' separate L2 caches); every level runs as its own launch");
}

// Installs `t`; a new chunk size or compression depth rebuilds the fused
// forward program (and the grouped compressions) from the device's renamed
// ops. Captured graphs are dropped (they bake in kernels). A new depth also
// drops a held forward pass: its reverse needs the groups it ran.
void ApplyTuning(Impl &w, const Tuning &t) {
  ValidateTuning(t);
  for (Impl::TreeGroup &g : w.groups) ApplyTuning(*g.w, t);
  const bool rebuild =
      !w.fused.empty() && (t.gather_chunk != w.tuning.gather_chunk || t.compress_depth != w.tuning.compress_depth);
  if (t.compress_depth != w.tuning.compress_depth || t.input_leaves != w.tuning.input_leaves)
    w.forward = Impl::Forward::kNone;
  w.tuning = t;
  for (Impl::GraphCache &g : w.graphs) g.Release();
  if (!rebuild) return;
  std::vector<DeviceRake> rakes(w.plan->rakes.size());
  if (!rakes.empty())
    Hip(hipMemcpy(rakes.data(), w.rakes.ptr, rakes.size() * sizeof(DeviceRake), hipMemcpyDeviceToHost),
        "GPU runtime Memcpy rake download");
  std::vector<DeviceCompress> compresses(w.plan->compressions.size());
  if (!compresses.empty())
    Hip(hipMemcpy(compresses.data(), w.compresses.ptr, compresses.size() * sizeof(DeviceCompress),
                  hipMemcpyDeviceToHost),
        "GPU runtime Memcpy compression download");
  BuildPrograms(w, *w.plan, rakes, compresses, w.capacity);
  if (w.fused.empty()) throw std::logic_error("the fused forward program could not be rebuilt");
}

// Binary16 edge outputs need 2-byte alignment.
void CheckEdgeOutput(const Impl &w, const void *edge_out) {
  if (w.half_edges() && (reinterpret_cast<std::uintptr_t>(edge_out) & 1u) != 0)
    throw std::invalid_argument("half-precision edge output must be 2-byte aligned");
}

void CheckDevice(const Impl &w, std::size_t batch, const void *nodes, const void *edges) {
  if (w.plan == nullptr) throw std::logic_error("accelerator Workspace::Reserve must precede this call");
  if (w.categorical) throw std::invalid_argument("device-resident calls take dense potentials");
  if (batch == 0 || batch > w.capacity)
    throw std::invalid_argument("device-resident batch exceeds the reserved capacity");
  if (nodes == nullptr || (w.E() != 0 && edges == nullptr)) throw std::invalid_argument("device input is null");
}

}  // namespace

void Workspace::LogPartitionDevice(const void *nodes, const void *edges, std::size_t batch, double *log_z,
                                   void *stream) {
  Impl &w = *impl_;
  CheckDevice(w, batch, nodes, edges);
  if (log_z == nullptr) throw std::invalid_argument("log-partition output is null");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  Impl::GraphKey key;
  key.args[0] = nodes, key.args[1] = edges, key.args[2] = log_z, key.batch = batch;
  const auto forward = [](Impl &v, hipStream_t s, const void *nodes, const void *edges, std::size_t batch,
                          double *log_z) {
    const Launch l = v.MakeLaunch(v.real, batch, s);
    LaunchInitNodes(l, v.real, nodes, v.nodes.ptr, v.node_scale.as<double>());
    const LinearBuffers buf = v.Linear(edges, nullptr);
    SumProductForward(v, l, buf);
    LaunchFinish(l, buf, log_z);
  };
  RunOrReplay(w, 0, key, ChooseStream(w, stream), [&](hipStream_t s) {
    if (w.groups.empty()) return forward(w, s, nodes, edges, batch, log_z);
    for (Impl::TreeGroup &g : w.groups)
      forward(*g.w, s, At(nodes, g.first_node, w.states, w.S()), At(edges, g.first_edge, w.KK(), w.S()), 1,
              log_z + g.first_root);
  });
}

void Workspace::MarginalsDevice(const void *nodes, const void *edges, std::size_t batch, double *log_z,
                                void *node_out, void *edge_out, void *stream, thmm_launch_timing *timings,
                                std::size_t capacity, std::size_t *count) {
  Impl &w = *impl_;
  CheckDevice(w, batch, nodes, edges);
  RequireMarginals(w);
  if (log_z == nullptr || node_out == nullptr || (w.E() != 0 && edge_out == nullptr))
    throw std::invalid_argument("marginal output is null");
  CheckEdgeOutput(w, edge_out);
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  const hipStream_t s = ChooseStream(w, stream);
  // With `timings` every launch is bracketed by an event pair (eager only).
  std::unique_ptr<Profiler> profiler;
  if (timings != nullptr) profiler = std::make_unique<Profiler>(s, static_cast<std::uint32_t>(batch));
  Profiler *p = profiler.get();
  const auto enqueue = [&](hipStream_t on) {
    EnqueueMarginalsAll(w, on, nodes, edges, batch, log_z, node_out, edge_out, p);
  };
  if (p) {
    enqueue(s);
    CheckLaunches();
    *count = p->Report(timings, capacity);
    CheckPersistent(w);
    return;
  }
  Impl::GraphKey key;
  key.args[0] = nodes, key.args[1] = edges, key.args[2] = log_z, key.args[3] = node_out, key.args[4] = edge_out;
  key.batch = batch;
  RunOrReplay(w, 1, key, s, enqueue);
}

void Workspace::MaximumDevice(const void *nodes, const void *edges, std::size_t batch, double *log_weights,
                              std::uint32_t *states, void *stream) {
  Impl &w = *impl_;
  CheckDevice(w, batch, nodes, edges);
  RequireMaximum(w);
  if (log_weights == nullptr || states == nullptr) throw std::invalid_argument("MAP output is null");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  RunMaximum(w, batch, Source{}, DenseCall{nodes, edges, nullptr, log_weights, states, nullptr, ChooseStream(w, stream)});
}

void Workspace::SampleDevice(const void *nodes, const void *edges, const void *uniforms, std::size_t batch,
                             double *log_z, std::uint32_t *states, std::uint32_t *failure, void *stream) {
  Impl &w = *impl_;
  CheckDevice(w, batch, nodes, edges);
  if (!(w.flags & THMM_RESERVE_SAMPLING))
    throw std::invalid_argument("prepared accelerator posterior sampling requires ReserveSampling");
  if (uniforms == nullptr || log_z == nullptr || states == nullptr || failure == nullptr)
    throw std::invalid_argument("sampling input or output is null");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  RunSample(w, batch, Source{}, DenseCall{nodes, edges, uniforms, log_z, states, failure, ChooseStream(w, stream)});
}

namespace {

// The log-potential pair's compressions read edge slots as log potentials
// in place (the K = 32 and 64 FP32 MFMA kernels) instead of staged.
bool CompressionsReadLogEdges(const Impl &w) { return w.real == Real::kF32 && (w.states == 32 || w.states == 64); }

// How a log-potential forward pass reads its edges: with the fused forward
// the team rake kernels read the log edges in place and only the edges
// compressions read are staged; at K = 32 and 64 (FP32) the MFMA
// compressions read theirs in place as well.
struct LogMode {
  bool in_place = false, compress_in_place = false;
};

LogMode LogModeOf(const Impl &w, const void *log_edges, std::size_t batch) {
  LogMode m;
  m.in_place = w.tuning.rake_log_edges != 0 && !w.fused.empty() &&
               static_cast<std::uint64_t>(w.max_rake_level) * batch * 64 < (1ull << 32) &&
               (reinterpret_cast<std::uintptr_t>(log_edges) & 15u) == 0;
  m.compress_in_place = m.in_place && CompressionsReadLogEdges(w);
  return m;
}

// The staged edge buffer, allocated when a device-resident log-potential
// call first needs it.
void EnsureEdgeStaging(Impl &w) {
  if (w.edges.ptr != nullptr || w.E() == 0) return;
  if (w.group_member)
    throw std::logic_error(
        "a workspace in tree groups reads log edges in place at K = 32 and 64: keep rake_log_edges = 1 and "
        "16-byte aligned log edges and gradients");
  w.edges.Allocate(Product({w.E(), w.KK(), w.S()}, "edge storage"), "edge storage");
}

// Stages what a log-potential forward pass reads of its edges and returns
// the buffers its levels use.
LinearBuffers LogEdgeBuffers(Impl &w, const Launch &l, const void *log_edges, const LogMode &m) {
  LinearBuffers buf = w.Linear(w.edges.ptr, w.edge_scale.as<double>());
  if (m.in_place && (m.compress_in_place ||
                     LaunchStageLogEdges(l, log_edges, w.compress_edges.as<std::uint32_t>(), w.compress_edge_count,
                                         w.edges.ptr, w.edge_scale.as<double>()))) {
    buf.log_edges = log_edges;
    buf.log_edge_shift = w.edge_scale.as<double>();  // rows of rake edges; staged rows keep their scales
    buf.compress_log_edges = m.compress_in_place ? 1 : 0;
  } else {
    LaunchInitEdgesFromLogs(l, log_edges, w.edges.ptr, w.edge_scale.as<double>());
  }
  return buf;
}

void EnqueueLogZForward(Impl &w, hipStream_t s, const void *log_nodes, const void *log_edges, std::size_t batch,
                        double *log_z, const LogMode &m) {
  const Launch l = w.MakeLaunch(w.real, batch, s);
  LaunchInitNodesFromLogs(l, log_nodes, w.nodes.ptr, w.node_scale.as<double>());
  const LinearBuffers buf = LogEdgeBuffers(w, l, log_edges, m);
  SumProductForward(w, l, buf);
  LaunchFinish(l, buf, log_z);
}

// Records what the backward half needs of a forward pass.
void HoldLogForward(Impl &w, const void *log_edges, std::size_t batch, const LogMode &m) {
  w.forward_log_edges =
      m.in_place && (w.states == 8 || w.states == 16 || w.states == 32 || w.states == 64) ? log_edges : nullptr;
  w.forward_log_source = log_edges;
  w.forward_in_place = m.in_place;
  w.forward_compress_in_place = m.compress_in_place;
  w.last_batch = batch;
}

// The compression levels of the forward pass that SumProductForward ran for
// these buffers. They rewrite every stored path from its inputs (edges,
// earlier paths and final node values), the values the forward pass wrote,
// bit for bit; rakes and absorptions are not repeated (they add into nodes).
void RecomputePaths(const Impl &w, const Launch &l, const LinearBuffers &buf) {
  if (UsesFusedForward(w, buf)) {
    for (const Impl::FusedStep &step : w.fused)
      if (step.kind == 3 || step.kind == kChainPrimitive) LaunchForwardLevel(l, step.kind, step.offset, step.count, buf);
  } else {
    for (const Level &level : w.levels)
      if (level.primitive == THMM_COMPRESSION) LaunchForwardLevel(l, level.primitive, level.offset, level.count, buf);
  }
}

bool Aligned16(const void *p) { return (reinterpret_cast<std::uintptr_t>(p) & 15u) == 0; }

// Whether the backward half stages every edge again (unaligned outputs keep
// the plan-level kernels, which read staged edges).
bool BackwardRestages(const Impl &w, const void *grad_nodes, const void *grad_edges, bool reduce_edges) {
  const void *edge_out = reduce_edges ? w.edge_out.ptr : grad_edges;
  const void *node_adj = w.last_batch == 1 ? grad_nodes : w.node_adj.ptr;
  return w.forward_log_edges != nullptr && !(Aligned16(edge_out) && Aligned16(node_adj));
}

void EnqueueLogZBackward(Impl &w, hipStream_t s, const double *grad, void *grad_nodes, void *grad_edges,
                         bool reduce_edges) {
  const Launch l = w.MakeLaunch(w.real, w.last_batch, s);
  LinearBuffers buf = w.Linear(w.edges.ptr, w.edge_scale.as<double>());
  buf.edge_out = reduce_edges ? w.edge_out.ptr : grad_edges;
  AdjointsInto(buf, w.last_batch, grad_nodes);
  if (w.forward_log_edges != nullptr) {
    // The team rake kernels need 16-byte aligned outputs; otherwise the
    // rest of the edges are staged for the plan-level kernels.
    if (Aligned16(buf.edge_out) && Aligned16(buf.node_adj)) {
      buf.log_edges = w.forward_log_edges;
      buf.log_edge_shift = w.edge_scale.as<double>();
      buf.compress_log_edges = CompressionsReadLogEdges(w) ? 1 : 0;
    } else
      LaunchInitEdgesFromLogs(l, w.forward_log_edges, w.edges.ptr, w.edge_scale.as<double>());
  }
  LaunchSeedAdjoint(l, buf, grad);
  RunReverse(w, l, buf);
  if (w.last_batch != 1) LaunchNodeAdjointOut(l, w.node_adj.ptr, grad_nodes);
  if (reduce_edges) LaunchReduceEdgesOverBatch(l, w.edge_out.ptr, grad_edges);
}

}  // namespace

void Workspace::LogZForwardDevice(const void *log_nodes, const void *log_edges, std::size_t batch, double *log_z,
                                  void *stream) {
  Impl &w = *impl_;
  CheckDevice(w, batch, log_nodes, log_edges);
  if (log_z == nullptr) throw std::invalid_argument("log-partition output is null");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  Impl::GraphKey key;
  key.args[0] = log_nodes, key.args[1] = log_edges, key.args[2] = log_z, key.batch = batch;
  const std::size_t K = w.states, S = w.S();
  if (!w.groups.empty()) {
    // Tree groups: each group's forward pass in turn; only node state (and
    // edge row shifts) outlive a group, the backward half recomputes paths.
    std::vector<LogMode> modes;
    for (Impl::TreeGroup &g : w.groups) {
      modes.push_back(LogModeOf(*g.w, At(log_edges, g.first_edge, K * K, S), 1));
      if (!modes.back().compress_in_place) EnsureEdgeStaging(*g.w);
    }
    RunOrReplay(w, 2, key, ChooseStream(w, stream), [&](hipStream_t s) {
      for (std::size_t i = 0; i < w.groups.size(); ++i) {
        const Impl::TreeGroup &g = w.groups[i];
        EnqueueLogZForward(*g.w, s, At(log_nodes, g.first_node, K, S), At(log_edges, g.first_edge, K * K, S), 1,
                           log_z + g.first_root, modes[i]);
      }
    });
    for (std::size_t i = 0; i < w.groups.size(); ++i)
      HoldLogForward(*w.groups[i].w, At(log_edges, w.groups[i].first_edge, K * K, S), 1, modes[i]);
    HoldLogForward(w, log_edges, batch, LogMode{});
    w.forward_log_edges = log_edges;  // names the forward pass in the backward half's graph key
    w.forward = Impl::Forward::kLogPotentials;
    return;
  }
  const LogMode mode = LogModeOf(w, log_edges, batch);
  if (!mode.compress_in_place) EnsureEdgeStaging(w);
  key.flag = mode.in_place ? 1 : 0;
  w.forward_log_edges = nullptr;
  RunOrReplay(w, 2, key, ChooseStream(w, stream), [&](hipStream_t s) {
    EnqueueLogZForward(w, s, log_nodes, log_edges, batch, log_z, mode);
  });
  HoldLogForward(w, log_edges, batch, mode);
  w.forward = Impl::Forward::kLogPotentials;
}

void Workspace::LogZBackwardDevice(const double *grad, void *grad_nodes, void *grad_edges, bool reduce_edges,
                                   void *stream) {
  Impl &w = *impl_;
  if (w.forward != Impl::Forward::kLogPotentials)
    throw std::logic_error("thmm_logz_backward_device requires a preceding thmm_logz_forward_device call");
  RequireMarginals(w);
  if (grad_nodes == nullptr || (w.E() != 0 && grad_edges == nullptr))
    throw std::invalid_argument("gradient output is null");
  CheckEdgeOutput(w, grad_edges);
  if (reduce_edges && w.E() != 0 && w.edge_out.ptr == nullptr)
    throw std::logic_error("batch-reduced edge gradients need workspace edge storage (no device-resident flag)");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  Impl::GraphKey key;
  key.args[0] = grad, key.args[1] = grad_nodes, key.args[2] = grad_edges, key.batch = w.last_batch;
  key.args[3] = w.forward_log_edges;
  key.flag = reduce_edges ? 1 : 0;
  const std::size_t K = w.states, S = w.S(), out_s = w.half_edges() ? 2 : S;
  if (!w.groups.empty()) {
    // Per group: stage its edges again as its forward pass did, recompute
    // its paths, then its reverse pass.
    for (Impl::TreeGroup &g : w.groups)
      if (BackwardRestages(*g.w, At(grad_nodes, g.first_node, K, S), At(grad_edges, g.first_edge, K * K, out_s),
                           false))
        EnsureEdgeStaging(*g.w);
    RunOrReplay(w, 3, key, ChooseStream(w, stream), [&](hipStream_t s) {
      for (Impl::TreeGroup &g : w.groups) {
        Impl &v = *g.w;
        const Launch l = v.MakeLaunch(v.real, 1, s);
        const LogMode mode{v.forward_in_place, v.forward_compress_in_place};
        RecomputePaths(v, l, LogEdgeBuffers(v, l, v.forward_log_source, mode));
        EnqueueLogZBackward(v, s, grad != nullptr ? grad + g.first_root : nullptr,
                            At(grad_nodes, g.first_node, K, S), At(grad_edges, g.first_edge, K * K, out_s), false);
      }
    });
    return;
  }
  if (BackwardRestages(w, grad_nodes, grad_edges, reduce_edges)) EnsureEdgeStaging(w);
  RunOrReplay(w, 3, key, ChooseStream(w, stream), [&](hipStream_t s) {
    EnqueueLogZBackward(w, s, grad, grad_nodes, grad_edges, reduce_edges);
  });
}

void Workspace::SetTuning(const Tuning &t) {
  Impl &w = *impl_;
  if (w.plan == nullptr) throw std::logic_error("accelerator Workspace::Reserve must precede this call");
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  Hip(hipStreamSynchronize(w.stream), "tree-HMM accelerator execution");
  ApplyTuning(w, t);
}

Tuning Workspace::GetTuning() const { return impl_->tuning; }

// Coordinate descent over the launch variants: each candidate runs the whole
// marginal pass on the caller's inputs (eagerly, on `stream`), timed with an
// event pair as the best of three runs; a candidate replaces the current
// value when it is at least 1 % faster. Outputs hold the last run's (valid)
// results.
double Workspace::AutotuneMarginalsDevice(const void *nodes, const void *edges, std::size_t batch, double *log_z,
                                          void *node_out, void *edge_out, void *stream) {
  Impl &w = *impl_;
  CheckDevice(w, batch, nodes, edges);
  RequireMarginals(w);
  if (log_z == nullptr || node_out == nullptr || (w.E() != 0 && edge_out == nullptr))
    throw std::invalid_argument("marginal output is null");
  CheckEdgeOutput(w, edge_out);
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  w.forward = Impl::Forward::kNone;
  const hipStream_t s = ChooseStream(w, stream);
  hipEvent_t start = nullptr, stop = nullptr;
  Hip(hipEventCreate(&start), "GPU runtime EventCreate");
  Hip(hipEventCreate(&stop), "GPU runtime EventCreate");
  const auto time_current = [&]() {
    float best = std::numeric_limits<float>::infinity();
    for (int rep = 0; rep < 3; ++rep) {
      Hip(hipEventRecord(start, s), "GPU runtime EventRecord");
      EnqueueMarginalsAll(w, s, nodes, edges, batch, log_z, node_out, edge_out, nullptr);
      Hip(hipEventRecord(stop, s), "GPU runtime EventRecord");
      Hip(hipEventSynchronize(stop), "tree-HMM accelerator execution");
      CheckLaunches();
      float ms = 0.0f;
      Hip(hipEventElapsedTime(&ms, start, stop), "GPU runtime EventElapsedTime");
      best = std::min(best, ms);
    }
    return static_cast<double>(best);
  };
  double best_ms = 0.0;
  try {
    ApplyTuning(w, w.tuning);
    best_ms = time_current();
    const auto search = [&](std::uint32_t Tuning::*knob, std::initializer_list<std::uint32_t> values) {
      const Tuning kept = w.tuning;
      Tuning best = kept;
      for (std::uint32_t v : values) {
        if (v == kept.*knob) continue;
        Tuning t = kept;
        t.*knob = v;
        ApplyTuning(w, t);
        const double ms = time_current();
        if (ms < best_ms * 0.99) best_ms = ms, best = t;
      }
      ApplyTuning(w, best);
    };
    const bool mfma16 = w.states == 16 && w.real == Real::kF32;
    if (!w.fused.empty()) search(&Tuning::gather_chunk, {4, 8, 16, 32});
    if (mfma16) search(&Tuning::compress_items_forward, {1, 2});
    if (mfma16) search(&Tuning::compress_items_reverse, {1, 2});
    if (mfma16 && !w.fused.empty() && !w.plan->compressions.empty()) search(&Tuning::compress_depth, {1, 2, 3});
    search(&Tuning::rake_block, {64, 128, 256});
    if (kPersistentPassesEnabled && w.persistent_forward.num_tasks != 0 && batch == 1)
      search(&Tuning::persistent, {0, 1});
  } catch (...) {
    static_cast<void>(hipEventDestroy(start));
    static_cast<void>(hipEventDestroy(stop));
    throw;
  }
  static_cast<void>(hipEventDestroy(start));
  static_cast<void>(hipEventDestroy(stop));
  return best_ms;
}

void Workspace::SetPathBudget(std::size_t bytes) { impl_->path_budget = bytes; }



void Workspace::PathStorage(std::size_t *groups, std::size_t *bytes) const {
  const Impl &w = *impl_;
  if (w.plan == nullptr) throw std::logic_error("accelerator Workspace::Reserve must precede this call");
  *groups = w.groups.size();
  *bytes = w.groups.empty() ? w.paths.bytes + w.path_scale.bytes + w.path_adj.bytes : [&] {
    const Impl &g = *w.groups.front().w;
    return g.paths.bytes + g.path_scale.bytes + g.path_adj.bytes;
  }();
}

void Workspace::Synchronize() {
  Impl &w = *impl_;
  if (w.stream == nullptr) return;
  Hip(hipSetDevice(w.device), "GPU runtime SetDevice");
  Hip(hipStreamSynchronize(w.stream), "tree-HMM accelerator execution");
  CheckPersistent(w);
}

}  // namespace thmm
