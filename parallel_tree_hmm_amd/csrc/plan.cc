// Rake-compress contraction planner; see plan.h for the level structure.
#include "plan.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <exception>
#include <limits>
#include <mutex>
#include <numeric>
#include <thread>
#include <utility>

namespace thmm {
namespace {

constexpr Index kNone = std::numeric_limits<Index>::max();

// Checks run once per node or op: the message is built only on failure.
void Require(bool condition, const char *message) {
  if (!condition) throw std::invalid_argument(message);
}
template <typename Message> void RequireWith(bool condition, Message &&message) {
  if (!condition) throw std::invalid_argument(message());
}

// Returns false when the parent links contain a cycle. parent[i] == kNone
// marks a root.
bool Acyclic(const std::vector<Index> &parent) {
  const std::size_t n = parent.size();
  std::vector<unsigned char> state(n, 0);  // 0 new, 1 on stack, 2 done
  std::vector<Index> path;
  for (std::size_t start = 0; start < n; ++start) {
    if (state[start] != 0) continue;
    path.clear();
    Index node = static_cast<Index>(start);
    while (node != kNone && state[node] == 0) {
      state[node] = 1;
      path.push_back(node);
      node = parent[node];
    }
    if (node != kNone && state[node] == 1) return false;
    for (Index visited : path) state[visited] = 2;
  }
  return true;
}

}  // namespace

thmm_plan_arrays Plan::Arrays() const {
  thmm_plan_arrays a{};
  a.num_nodes = num_nodes;
  a.num_edges = num_edges;
  a.num_branches = num_branches;
  a.num_compressions = compressions.size();
  a.num_roots = roots.size();
  a.num_rakes = rakes.size();
  a.num_combinations = combinations.size();
  a.num_absorptions = absorptions.size();
  a.num_batches = batches.size();
  a.num_rounds = num_rounds();
  a.roots = roots.data();
  a.edge_parents = edge_parents.data();
  a.edge_children = edge_children.data();
  a.rakes = rakes.data();
  a.combinations = combinations.data();
  a.absorptions = absorptions.data();
  a.compressions = compressions.data();
  a.batches = batches.data();
  a.round_batch_offsets = round_batch_offsets.data();
  return a;
}

namespace {

// Contracts the forest `parent` (kNone marks a root) round by round into
// `plan`: edge arrays, ops, batches and round offsets. Ops are listed in
// ascending node order within each batch, so the plan of a forest whose
// trees occupy consecutive node ranges is, round by round and batch by
// batch, the concatenation of its trees' plans.
void Contract(const std::vector<Index> &parent, Plan &plan) {
  const std::size_t n = parent.size();
  // Canonical edge order: one edge per non-root node, in ascending node
  // order; edge e joins edge_parents[e] -> edge_children[e].
  std::vector<Index> slot(n, kNone);  // slot joining cur_parent[i] -> i
  for (std::size_t i = 0; i < n; ++i) {
    if (parent[i] == kNone) continue;
    slot[i] = static_cast<Index>(plan.edge_parents.size());
    plan.edge_parents.push_back(parent[i]);
    plan.edge_children.push_back(static_cast<Index>(i));
  }
  plan.num_edges = plan.edge_parents.size();

  std::vector<Index> cur_parent = parent;
  std::vector<Index> child_count(n, 0);
  std::vector<Index> child_xor(n, 0);  // XOR of live children ids
  for (std::size_t i = 0; i < n; ++i) {
    if (parent[i] == kNone) continue;
    ++child_count[parent[i]];
    child_xor[parent[i]] ^= static_cast<Index>(i);
  }
  std::vector<unsigned char> alive(n, 1);
  std::vector<Index> live;  // live non-root nodes, ascending
  live.reserve(n);
  for (std::size_t i = 0; i < n; ++i)
    if (parent[i] != kNone) live.push_back(static_cast<Index>(i));

  std::vector<std::pair<Index, Index>> parent_branch;  // (parent, branch)
  std::vector<Index> group_start, group_size, work, selected;
  std::vector<unsigned char> unary(n, 0);
  Index round = 0;

  auto push_batch = [&](thmm_primitive primitive, std::size_t offset, std::size_t count) {
    if (count == 0) return;
    plan.batches.push_back({static_cast<std::uint32_t>(primitive),
                            static_cast<std::uint32_t>(offset),
                            static_cast<std::uint32_t>(count)});
  };

  while (!live.empty()) {
    plan.round_batch_offsets.push_back(static_cast<Index>(plan.batches.size()));

    // Rakes: every live leaf.
    const std::size_t rake_offset = plan.rakes.size();
    parent_branch.clear();
    for (Index leaf : live) {
      if (child_count[leaf] != 0) continue;
      const Index branch = static_cast<Index>(plan.num_branches++);
      plan.rakes.push_back({leaf, cur_parent[leaf], slot[leaf], branch});
      parent_branch.emplace_back(cur_parent[leaf], branch);
    }
    push_batch(THMM_RAKE, rake_offset, plan.rakes.size() - rake_offset);
    for (std::size_t r = rake_offset; r < plan.rakes.size(); ++r) {
      const thmm_rake &rake = plan.rakes[r];
      --child_count[rake.parent];
      child_xor[rake.parent] ^= rake.leaf;
      alive[rake.leaf] = 0;
    }

    // Group the round's branches by parent (stable: branch order).
    std::stable_sort(parent_branch.begin(), parent_branch.end(),
                     [](const auto &a, const auto &b) { return a.first < b.first; });
    group_start.clear();
    group_size.clear();
    work.clear();
    for (std::size_t i = 0; i < parent_branch.size(); ++i) {
      if (i == 0 || parent_branch[i].first != parent_branch[i - 1].first) {
        group_start.push_back(static_cast<Index>(i));
        group_size.push_back(0);
      }
      ++group_size.back();
      work.push_back(parent_branch[i].second);
    }

    // Sibling reduction: pairwise stages until each parent has one branch.
    for (Index stage = 0;; ++stage) {
      const std::size_t offset = plan.combinations.size();
      for (std::size_t g = 0; g < group_start.size(); ++g) {
        const Index start = group_start[g];
        const Index size = group_size[g];
        if (size < 2) continue;
        Index kept = 0;
        for (Index j = 0; j < size; j += 2) {
          if (j + 1 < size)
            plan.combinations.push_back({work[start + j], work[start + j + 1], stage});
          work[start + kept++] = work[start + j];
        }
        group_size[g] = kept;
      }
      if (plan.combinations.size() == offset) break;
      push_batch(THMM_BRANCH_COMBINATION, offset, plan.combinations.size() - offset);
    }

    // Absorptions.
    const std::size_t absorb_offset = plan.absorptions.size();
    for (std::size_t g = 0; g < group_start.size(); ++g)
      plan.absorptions.push_back(
          {parent_branch[group_start[g]].first, work[group_start[g]], round});
    push_batch(THMM_BRANCH_ABSORPTION, absorb_offset, plan.absorptions.size() - absorb_offset);

    // Compressions: alternate nodes of maximal unary chains, chain heads
    // first.
    for (Index node : live)
      unary[node] = alive[node] && child_count[node] == 1 ? 1 : 0;
    selected.clear();
    for (Index node : live) {
      if (!unary[node]) continue;
      const Index up = cur_parent[node];
      if (up != kNone && unary[up]) continue;  // not a chain head
      Index x = node;
      for (std::size_t position = 0; unary[x]; ++position) {
        if (position % 2 == 0) selected.push_back(x);
        x = child_xor[x];
      }
    }
    const std::size_t compress_offset = plan.compressions.size();
    for (Index middle : selected) {
      const Index child = child_xor[middle];
      plan.compressions.push_back({slot[middle], middle, slot[child], cur_parent[middle], child,
                                   static_cast<Index>(plan.compressions.size())});
    }
    for (std::size_t c = compress_offset; c < plan.compressions.size(); ++c) {
      const thmm_compression &op = plan.compressions[c];
      child_xor[op.parent] ^= op.middle ^ op.child;
      cur_parent[op.child] = op.parent;
      slot[op.child] = op.left_edge;
      alive[op.middle] = 0;
    }
    push_batch(THMM_COMPRESSION, compress_offset, plan.compressions.size() - compress_offset);
    for (Index node : live) unary[node] = 0;

    live.erase(std::remove_if(live.begin(), live.end(), [&](Index i) { return !alive[i]; }),
               live.end());
    ++round;
  }
  plan.round_batch_offsets.push_back(static_cast<Index>(plan.batches.size()));
}

// Threads for planning a forest: THMM_PLAN_THREADS when set (1 plans
// sequentially), else the host's hardware threads, at most 16.
unsigned PlannerThreads() {
  if (const char *v = std::getenv("THMM_PLAN_THREADS")) {
    const long t = std::strtol(v, nullptr, 10);
    return t >= 1 ? static_cast<unsigned>(std::min<long>(t, 64)) : 1u;
  }
  return std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

// True when every tree occupies the node range from its root to the next
// root (roots ascending, node 0 a root): the trees are then planned apart.
bool ConsecutiveTrees(const std::vector<Index> &parent, const std::vector<Index> &roots) {
  if (roots.empty() || roots[0] != 0) return false;
  for (std::size_t t = 0; t < roots.size(); ++t) {
    const Index first = roots[t];
    const Index end = t + 1 < roots.size() ? roots[t + 1] : static_cast<Index>(parent.size());
    for (Index i = first + 1; i < end; ++i)
      if (parent[i] == kNone || parent[i] < first || parent[i] >= end) return false;
  }
  return true;
}

// Local op ranges of one tree in one round.
struct TreeRound {
  std::uint32_t rake_first = 0, rakes = 0, absorb_first = 0, absorbs = 0, compress_first = 0, compresses = 0;
  std::vector<std::pair<std::uint32_t, std::uint32_t>> stages;  // (first, count) per combination stage
};

std::vector<TreeRound> RoundsOf(const Plan &p) {
  std::vector<TreeRound> rounds(p.num_rounds());
  for (std::size_t r = 0; r < rounds.size(); ++r) {
    TreeRound &tr = rounds[r];
    bool seen_rake = false, seen_absorb = false, seen_compress = false;
    for (Index b = p.round_batch_offsets[r]; b < p.round_batch_offsets[r + 1]; ++b) {
      const thmm_primitive_batch &batch = p.batches[b];
      switch (batch.primitive) {
        case THMM_RAKE: tr.rake_first = batch.offset, tr.rakes = batch.count, seen_rake = true; break;
        case THMM_BRANCH_COMBINATION: tr.stages.push_back({batch.offset, batch.count}); break;
        case THMM_BRANCH_ABSORPTION:
          tr.absorb_first = batch.offset, tr.absorbs = batch.count, seen_absorb = true;
          break;
        default: tr.compress_first = batch.offset, tr.compresses = batch.count, seen_compress = true; break;
      }
    }
    // Empty batches are not listed; their ranges start where the ops of
    // that kind continue.
    if (!seen_rake) tr.rake_first = r == 0 ? 0 : rounds[r - 1].rake_first + rounds[r - 1].rakes;
    if (!seen_absorb) tr.absorb_first = r == 0 ? 0 : rounds[r - 1].absorb_first + rounds[r - 1].absorbs;
    if (!seen_compress) tr.compress_first = r == 0 ? 0 : rounds[r - 1].compress_first + rounds[r - 1].compresses;
  }
  return rounds;
}

// Plans each tree of a consecutive forest on its own thread pool slot and
// merges the plans round by round. Branch and compression ids are numbered
// round-major and tree-minor, as the single contraction numbers them, so
// the result equals Contract() over the whole forest.
void ContractTrees(const std::vector<Index> &parent, Plan &plan, unsigned threads) {
  const std::vector<Index> &roots = plan.roots;
  const std::size_t trees = roots.size(), n = parent.size();
  const auto first = [&](std::size_t t) { return roots[t]; };
  const auto end = [&](std::size_t t) { return t + 1 < trees ? roots[t + 1] : static_cast<Index>(n); };
  std::vector<Plan> local(trees);
  std::vector<std::vector<TreeRound>> rounds(trees);
  std::vector<std::size_t> order(trees);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
    return end(a) - first(a) > end(b) - first(b);
  });
  const auto parallel = [&](auto &&job) {
    std::atomic<std::size_t> next{0};
    std::exception_ptr failure;
    std::mutex failure_lock;
    const auto work = [&] {
      for (std::size_t i; (i = next.fetch_add(1)) < trees;) {
        try {
          job(order[i]);
        } catch (...) {
          std::lock_guard<std::mutex> hold(failure_lock);
          if (!failure) failure = std::current_exception();
        }
      }
    };
    std::vector<std::thread> pool;
    for (unsigned k = 1; k < std::min<std::size_t>(threads, trees); ++k) pool.emplace_back(work);
    work();
    for (std::thread &th : pool) th.join();
    if (failure) std::rethrow_exception(failure);
  };
  parallel([&](std::size_t t) {
    std::vector<Index> sub(parent.begin() + first(t), parent.begin() + end(t));
    for (Index &p : sub) p = p == kNone ? kNone : p - first(t);
    Contract(sub, local[t]);
    rounds[t] = RoundsOf(local[t]);
  });

  std::size_t round_count = 0, max_stages = 0;
  for (std::size_t t = 0; t < trees; ++t) {
    round_count = std::max(round_count, rounds[t].size());
    for (const TreeRound &tr : rounds[t]) max_stages = std::max(max_stages, tr.stages.size());
  }
  // Global offsets per (round, tree) of each op kind, round-major.
  const auto at = [&](std::size_t r, std::size_t t) { return r * trees + t; };
  const std::size_t cells = round_count * trees;
  std::vector<std::size_t> rake_at(cells), absorb_at(cells), compress_at(cells), stage_at(cells * max_stages);
  std::vector<thmm_primitive_batch> batches;
  std::vector<Index> round_offsets;
  std::size_t rakes = 0, combinations = 0, absorptions = 0, compressions = 0;
  const auto round_of = [&](std::size_t t, std::size_t r) -> const TreeRound * {
    return r < rounds[t].size() ? &rounds[t][r] : nullptr;
  };
  for (std::size_t r = 0; r < round_count; ++r) {
    round_offsets.push_back(static_cast<Index>(batches.size()));
    const auto push = [&](thmm_primitive primitive, std::size_t offset, std::size_t count) {
      if (count != 0)
        batches.push_back({static_cast<std::uint32_t>(primitive), static_cast<std::uint32_t>(offset),
                           static_cast<std::uint32_t>(count)});
    };
    std::size_t begin = rakes;
    for (std::size_t t = 0; t < trees; ++t)
      if (const TreeRound *tr = round_of(t, r)) rake_at[at(r, t)] = rakes, rakes += tr->rakes;
    push(THMM_RAKE, begin, rakes - begin);
    for (std::size_t s = 0; s < max_stages; ++s) {
      begin = combinations;
      for (std::size_t t = 0; t < trees; ++t) {
        const TreeRound *tr = round_of(t, r);
        if (tr != nullptr && s < tr->stages.size())
          stage_at[at(r, t) * max_stages + s] = combinations, combinations += tr->stages[s].second;
      }
      push(THMM_BRANCH_COMBINATION, begin, combinations - begin);
    }
    begin = absorptions;
    for (std::size_t t = 0; t < trees; ++t)
      if (const TreeRound *tr = round_of(t, r)) absorb_at[at(r, t)] = absorptions, absorptions += tr->absorbs;
    push(THMM_BRANCH_ABSORPTION, begin, absorptions - begin);
    begin = compressions;
    for (std::size_t t = 0; t < trees; ++t)
      if (const TreeRound *tr = round_of(t, r)) compress_at[at(r, t)] = compressions, compressions += tr->compresses;
    push(THMM_COMPRESSION, begin, compressions - begin);
  }
  round_offsets.push_back(static_cast<Index>(batches.size()));

  // Edge e of tree t is global edge e + (first(t) - t): one root per tree.
  std::vector<std::size_t> edge_base(trees);
  for (std::size_t t = 0; t < trees; ++t) edge_base[t] = first(t) - t;
  plan.num_edges = n - trees;
  plan.num_branches = rakes;
  plan.edge_parents.resize(plan.num_edges);
  plan.edge_children.resize(plan.num_edges);
  plan.rakes.resize(rakes);
  plan.combinations.resize(combinations);
  plan.absorptions.resize(absorptions);
  plan.compressions.resize(compressions);
  parallel([&](std::size_t t) {
    const Plan &p = local[t];
    const Index node = first(t);
    const Index edge = static_cast<Index>(edge_base[t]);
    for (std::size_t e = 0; e < p.num_edges; ++e) {
      plan.edge_parents[edge + e] = p.edge_parents[e] + node;
      plan.edge_children[edge + e] = p.edge_children[e] + node;
    }
    for (std::size_t r = 0; r < rounds[t].size(); ++r) {
      const TreeRound &tr = rounds[t][r];
      // A round's branches are created by its rakes and consumed within it.
      const auto branch = [&](Index b) { return static_cast<Index>(b - tr.rake_first + rake_at[at(r, t)]); };
      for (std::uint32_t k = 0; k < tr.rakes; ++k) {
        const thmm_rake &op = p.rakes[tr.rake_first + k];
        plan.rakes[rake_at[at(r, t)] + k] = {op.leaf + node, op.parent + node, op.edge + edge, branch(op.branch)};
      }
      for (std::size_t s = 0; s < tr.stages.size(); ++s)
        for (std::uint32_t k = 0; k < tr.stages[s].second; ++k) {
          const thmm_branch_combination &op = p.combinations[tr.stages[s].first + k];
          plan.combinations[stage_at[at(r, t) * max_stages + s] + k] = {branch(op.destination), branch(op.source),
                                                                        op.stage};
        }
      for (std::uint32_t k = 0; k < tr.absorbs; ++k) {
        const thmm_branch_absorption &op = p.absorptions[tr.absorb_first + k];
        plan.absorptions[absorb_at[at(r, t)] + k] = {op.parent + node, branch(op.branch), op.round};
      }
      for (std::uint32_t k = 0; k < tr.compresses; ++k) {
        const thmm_compression &op = p.compressions[tr.compress_first + k];
        const Index index = static_cast<Index>(compress_at[at(r, t)] + k);
        plan.compressions[index] = {op.left_edge + edge, op.middle + node, op.right_edge + edge,
                                    op.parent + node,    op.child + node,  index};
      }
    }
  });
  plan.batches = std::move(batches);
  plan.round_batch_offsets = std::move(round_offsets);
}

}  // namespace

Plan MakePlan(const std::int64_t *parents, std::size_t num_nodes) {
  if (num_nodes == 0) throw std::invalid_argument("a plan needs at least one node");
  if (num_nodes >= kNone) throw std::length_error("plan node count exceeds uint32_t");
  const std::size_t n = num_nodes;
  std::vector<Index> parent(n, kNone);
  Plan plan;
  plan.num_nodes = n;
  for (std::size_t i = 0; i < n; ++i) {
    const std::int64_t p = parents[i];
    if (p == -1) {
      plan.roots.push_back(static_cast<Index>(i));
      continue;
    }
    RequireWith(p >= 0 && static_cast<std::size_t>(p) < n && static_cast<std::size_t>(p) != i,
                [&] { return "parent of node " + std::to_string(i) + " is out of range"; });
    parent[i] = static_cast<Index>(p);
  }
  Require(!plan.roots.empty(), "a plan needs at least one root (parent -1)");
  Require(Acyclic(parent), "parent links contain a cycle");
  // Forests of many trees in consecutive node ranges plan tree by tree in
  // parallel (same plan as the single contraction).
  constexpr std::size_t kParallelNodes = 1u << 16;
  const unsigned threads = PlannerThreads();
  if (threads > 1 && n >= kParallelNodes && plan.roots.size() > 1 && ConsecutiveTrees(parent, plan.roots))
    ContractTrees(parent, plan, threads);
  else
    Contract(parent, plan);
  return plan;
}

void Validate(const Plan &plan) {
  const std::size_t n = plan.num_nodes;
  const std::size_t e_count = plan.num_edges;
  Require(n > 0 && n < kNone, "plan node count is invalid");
  Require(plan.edge_parents.size() == e_count && plan.edge_children.size() == e_count,
          "edge arrays do not match num_edges");
  std::vector<Index> parent(n, kNone);
  std::vector<unsigned char> is_root(n, 0);
  for (Index r : plan.roots) {
    Require(r < n && !is_root[r], "roots must be distinct valid nodes");
    is_root[r] = 1;
  }
  for (std::size_t e = 0; e < e_count; ++e) {
    const Index p = plan.edge_parents[e], c = plan.edge_children[e];
    RequireWith(p < n && c < n && p != c, [&] { return "edge " + std::to_string(e) + " has invalid endpoints"; });
    RequireWith(!is_root[c] && parent[c] == kNone,
                [&] { return "node " + std::to_string(c) + " is the child of zero or several edges"; });
    parent[c] = p;
  }
  Require(e_count + plan.roots.size() == n, "every non-root node needs exactly one edge");
  Require(Acyclic(parent), "edges contain a cycle");
  Require(plan.rakes.size() + plan.compressions.size() == e_count,
          "rakes + compressions must eliminate every non-root node once");

  // Replay state.
  std::vector<Index> slot_parent(plan.edge_parents), slot_child(plan.edge_children);
  std::vector<unsigned char> slot_live(e_count, 1), alive(n, 1);
  std::vector<Index> child_count(n, 0), pending(n, 0);
  for (std::size_t e = 0; e < e_count; ++e) ++child_count[plan.edge_parents[e]];
  std::vector<Index> branch_target(plan.num_branches, kNone);
  std::vector<unsigned char> branch_state(plan.num_branches, 0);  // 0 new 1 live 2 used
  std::vector<unsigned char> tape_used(plan.compressions.size(), 0);
  std::vector<unsigned char> mark(std::max(n, std::max(e_count, plan.num_branches)), 0);
  std::vector<Index> touched;

  auto clear_marks = [&] {
    for (Index t : touched) mark[t] = 0;
    touched.clear();
  };
  auto claim = [&](Index id, const char *what) {
    RequireWith(!mark[id],
                [&] { return std::string("batch is not write-disjoint: ") + what + " " + std::to_string(id); });
    mark[id] = 1;
    touched.push_back(id);
  };

  std::size_t seen_rakes = 0, seen_combinations = 0, seen_absorptions = 0, seen_compressions = 0;
  for (const thmm_primitive_batch &batch : plan.batches) {
    Require(batch.count > 0, "empty primitive batch");
    switch (batch.primitive) {
      case THMM_RAKE: {
        Require(batch.offset == seen_rakes && batch.offset + batch.count <= plan.rakes.size(),
                "rake batches must tile the rake array in order");
        seen_rakes += batch.count;
        for (std::size_t i = batch.offset; i < batch.offset + batch.count; ++i) {
          const thmm_rake &op = plan.rakes[i];
          Require(op.leaf < n && alive[op.leaf] && !is_root[op.leaf], "rake leaf is not live");
          Require(child_count[op.leaf] == 0 && pending[op.leaf] == 0,
                  "rake leaf still has live children or unabsorbed branches");
          Require(op.edge < e_count && slot_live[op.edge] && slot_parent[op.edge] == op.parent &&
                      slot_child[op.edge] == op.leaf,
                  "rake edge slot does not join parent and leaf");
          Require(op.branch < plan.num_branches && branch_state[op.branch] == 0,
                  "rake branch is reused");
          claim(op.leaf, "node");
        }
        clear_marks();
        for (std::size_t i = batch.offset; i < batch.offset + batch.count; ++i) {
          const thmm_rake &op = plan.rakes[i];
          alive[op.leaf] = 0;
          slot_live[op.edge] = 0;
          --child_count[op.parent];
          ++pending[op.parent];
          branch_state[op.branch] = 1;
          branch_target[op.branch] = op.parent;
        }
        break;
      }
      case THMM_BRANCH_COMBINATION: {
        Require(batch.offset == seen_combinations &&
                    batch.offset + batch.count <= plan.combinations.size(),
                "combination batches must tile the combination array in order");
        seen_combinations += batch.count;
        for (std::size_t i = batch.offset; i < batch.offset + batch.count; ++i) {
          const thmm_branch_combination &op = plan.combinations[i];
          Require(op.destination < plan.num_branches && op.source < plan.num_branches &&
                      branch_state[op.destination] == 1 && branch_state[op.source] == 1 &&
                      op.destination != op.source,
                  "combination operands are not live branches");
          Require(branch_target[op.destination] == branch_target[op.source],
                  "combined branches belong to different parents");
          claim(op.destination, "branch");
          claim(op.source, "branch");
        }
        clear_marks();
        for (std::size_t i = batch.offset; i < batch.offset + batch.count; ++i) {
          const thmm_branch_combination &op = plan.combinations[i];
          branch_state[op.source] = 2;
          --pending[branch_target[op.source]];
        }
        break;
      }
      case THMM_BRANCH_ABSORPTION: {
        Require(batch.offset == seen_absorptions &&
                    batch.offset + batch.count <= plan.absorptions.size(),
                "absorption batches must tile the absorption array in order");
        seen_absorptions += batch.count;
        for (std::size_t i = batch.offset; i < batch.offset + batch.count; ++i) {
          const thmm_branch_absorption &op = plan.absorptions[i];
          Require(op.branch < plan.num_branches && branch_state[op.branch] == 1 &&
                      op.parent < n && branch_target[op.branch] == op.parent && alive[op.parent],
                  "absorbed branch does not belong to a live parent");
          claim(op.parent, "node");
        }
        clear_marks();
        for (std::size_t i = batch.offset; i < batch.offset + batch.count; ++i) {
          const thmm_branch_absorption &op = plan.absorptions[i];
          branch_state[op.branch] = 2;
          --pending[op.parent];
        }
        break;
      }
      case THMM_COMPRESSION: {
        Require(batch.offset == seen_compressions &&
                    batch.offset + batch.count <= plan.compressions.size(),
                "compression batches must tile the compression array in order");
        seen_compressions += batch.count;
        for (std::size_t i = batch.offset; i < batch.offset + batch.count; ++i) {
          const thmm_compression &op = plan.compressions[i];
          Require(op.middle < n && alive[op.middle] && !is_root[op.middle] &&
                      child_count[op.middle] == 1 && pending[op.middle] == 0,
                  "compression middle is not a settled unary node");
          Require(op.left_edge < e_count && slot_live[op.left_edge] &&
                      slot_parent[op.left_edge] == op.parent &&
                      slot_child[op.left_edge] == op.middle,
                  "compression left slot does not join parent and middle");
          Require(op.right_edge < e_count && slot_live[op.right_edge] &&
                      slot_parent[op.right_edge] == op.middle &&
                      slot_child[op.right_edge] == op.child,
                  "compression right slot does not join middle and child");
          Require(op.tape < plan.compressions.size() && !tape_used[op.tape],
                  "compression tape is reused");
          tape_used[op.tape] = 1;
          claim(op.middle, "node");
        }
        // Independence: no middle is another op's parent or child.
        for (std::size_t i = batch.offset; i < batch.offset + batch.count; ++i) {
          const thmm_compression &op = plan.compressions[i];
          Require(!mark[op.parent] && !mark[op.child],
                  "compressions in one batch are adjacent");
        }
        clear_marks();
        for (std::size_t i = batch.offset; i < batch.offset + batch.count; ++i) {
          const thmm_compression &op = plan.compressions[i];
          alive[op.middle] = 0;
          slot_child[op.left_edge] = op.child;
          slot_live[op.right_edge] = 0;
          --child_count[op.middle];
        }
        break;
      }
      default:
        Require(false, "unknown primitive in plan");
    }
  }
  Require(seen_rakes == plan.rakes.size() && seen_combinations == plan.combinations.size() &&
              seen_absorptions == plan.absorptions.size() &&
              seen_compressions == plan.compressions.size(),
          "batches do not cover every op");
  for (std::size_t i = 0; i < n; ++i) {
    if (is_root[i]) {
      Require(alive[i] && child_count[i] == 0 && pending[i] == 0, "root is not fully contracted");
    } else {
      RequireWith(!alive[i], [&] { return "node " + std::to_string(i) + " is never eliminated"; });
    }
  }
  for (std::size_t b = 0; b < plan.num_branches; ++b)
    RequireWith(branch_state[b] == 2, [&] { return "branch " + std::to_string(b) + " is never consumed"; });
  Require(!plan.round_batch_offsets.empty() && plan.round_batch_offsets.front() == 0 &&
              plan.round_batch_offsets.back() == plan.batches.size() &&
              std::is_sorted(plan.round_batch_offsets.begin(), plan.round_batch_offsets.end()),
          "round offsets do not partition the batches");
}

// Replays the plan's edge slots: a compression's composed path lives in slot
// E + c, and later ops that name its left edge read that slot. Renaming
// happens after each whole batch: ops of one batch never read a slot another
// op of the same batch writes.
RenamedOps RenameSlots(const Plan &plan) {
  const std::size_t E = plan.num_edges;
  std::vector<Index> current(E);
  for (std::size_t e = 0; e < E; ++e) current[e] = static_cast<Index>(e);
  RenamedOps out;
  out.rakes.resize(plan.rakes.size());
  out.combines.resize(plan.combinations.size());
  out.absorbs.resize(plan.absorptions.size());
  out.compresses.resize(plan.compressions.size());
  for (const thmm_primitive_batch &b : plan.batches) {
    for (std::uint32_t i = b.offset; i < b.offset + b.count; ++i) {
      switch (b.primitive) {
        case THMM_RAKE: {
          const thmm_rake &op = plan.rakes[i];
          out.rakes[i] = {op.leaf, current[op.edge], op.branch, op.parent};
          break;
        }
        case THMM_BRANCH_COMBINATION:
          out.combines[i] = {plan.combinations[i].destination, plan.combinations[i].source};
          break;
        case THMM_BRANCH_ABSORPTION:
          out.absorbs[i] = {plan.absorptions[i].parent, plan.absorptions[i].branch};
          break;
        case THMM_COMPRESSION: {
          const thmm_compression &op = plan.compressions[i];
          out.compresses[i] = {current[op.left_edge], op.middle, current[op.right_edge], static_cast<Index>(E + i),
                               op.parent, op.child, 0, 0};
          break;
        }
      }
    }
    if (b.primitive == THMM_COMPRESSION)
      for (std::uint32_t i = b.offset; i < b.offset + b.count; ++i)
        current[plan.compressions[i].left_edge] = static_cast<Index>(E + i);
  }
  return out;
}

Plan PlanFromArrays(const thmm_plan_arrays &a) {
  Plan plan;
  if (a.num_nodes == 0 || a.num_nodes >= kNone)
    throw std::invalid_argument("plan node count is invalid");
  auto copy = [](auto *data, std::size_t count, auto &out) {
    if (count != 0 && data == nullptr) throw std::invalid_argument("plan array is null");
    out.assign(data, data + count);
  };
  plan.num_nodes = a.num_nodes;
  plan.num_edges = a.num_edges;
  plan.num_branches = a.num_branches;
  copy(a.roots, a.num_roots, plan.roots);
  copy(a.edge_parents, a.num_edges, plan.edge_parents);
  copy(a.edge_children, a.num_edges, plan.edge_children);
  copy(a.rakes, a.num_rakes, plan.rakes);
  copy(a.combinations, a.num_combinations, plan.combinations);
  copy(a.absorptions, a.num_absorptions, plan.absorptions);
  copy(a.compressions, a.num_compressions, plan.compressions);
  copy(a.batches, a.num_batches, plan.batches);
  if (a.round_batch_offsets != nullptr && a.num_rounds != 0) {
    plan.round_batch_offsets.assign(a.round_batch_offsets,
                                    a.round_batch_offsets + a.num_rounds + 1);
  } else {
    plan.round_batch_offsets = {0, static_cast<Index>(plan.batches.size())};
  }
  Validate(plan);
  return plan;
}

}  // namespace thmm
