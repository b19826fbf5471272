// FP64 CPU oracle (test infrastructure; see oracle.h). Restates the
// dispatchers of src/inference.cc over plan arrays. Edge fields follow the
// reference convention: a compression overwrites its left_edge slot.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <thread>
#include <vector>

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

struct DomainError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct InvalidArgument : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// src/inference.cc:95-98
double PotentialLog(double x) { return x > 0.0 ? std::log(x) : -kInf; }

// src/inference.cc:100-110: max-shifted log-sum-exp; all -inf gives -inf.
template <class F> double Lse(std::size_t size, F term) {
  double top = -kInf;
  for (std::size_t i = 0; i < size; ++i) top = std::max(top, term(i));
  if (!std::isfinite(top)) return top;
  double sum = 0.0;
  for (std::size_t i = 0; i < size; ++i) sum += std::exp(term(i) - top);
  return top + std::log(sum);
}

// src/inference.cc:112-126: strict '>' so the lowest index wins ties.
template <class F> std::size_t ArgMax(std::size_t size, F term, double *best) {
  std::size_t arg = 0;
  double value = term(0);
  for (std::size_t i = 1; i < size; ++i) {
    const double candidate = term(i);
    if (candidate > value) {
      arg = i;
      value = candidate;
    }
  }
  *best = value;
  return arg;
}

// src/inference.cc:128-146: ascending inverse CDF over exp(lw - normalizer),
// skipping unsupported states, falling back to the last supported one.
// `distance` receives min |u - cumulative| over the cumulative values the
// scan produced (the interior CDF boundaries it compared against).
template <class F>
std::size_t SampleLog(std::size_t size, F term, double normalizer, double u, double *distance) {
  double cumulative = 0.0;
  std::size_t last = size;
  double nearest = kInf;
  for (std::size_t i = 0; i < size; ++i) {
    const double lw = term(i);
    if (!std::isfinite(lw)) continue;
    last = i;
    cumulative += std::exp(lw - normalizer);
    nearest = std::min(nearest, std::fabs(u - cumulative));
    if (u < cumulative) {
      if (distance) *distance = nearest;
      return i;
    }
  }
  if (last == size) throw DomainError("cannot sample from an empty conditional distribution");
  if (distance) *distance = nearest;
  return last;
}

struct View {
  const thmm_plan_arrays &p;
  std::size_t k;
  std::size_t kk;
  View(const thmm_plan_arrays &plan, std::size_t states) : p(plan), k(states), kk(states * states) {
    if (states == 0) throw InvalidArgument("a tree HMM must have at least one state");
  }
};

// Visits batches in contraction order (btrc::Contract as driven by
// src/inference.cc:888,901,912,923,934).
template <class D> void Contract(const thmm_plan_arrays &p, D &d) {
  for (std::size_t b = 0; b < p.num_batches; ++b) {
    const thmm_primitive_batch &batch = p.batches[b];
    switch (batch.primitive) {
      case THMM_RAKE: d.Rake(p.rakes + batch.offset, batch.count); break;
      case THMM_BRANCH_COMBINATION: d.Combine(p.combinations + batch.offset, batch.count); break;
      case THMM_BRANCH_ABSORPTION: d.Absorb(p.absorptions + batch.offset, batch.count); break;
      case THMM_COMPRESSION: d.Compress(p.compressions + batch.offset, batch.count); break;
      default: throw InvalidArgument("unknown primitive");
    }
  }
}

// Reverse contraction: batches and ops in reverse (src/inference.cc:416-518).
template <class D> void Reverse(const thmm_plan_arrays &p, D &d) {
  for (std::size_t b = p.num_batches; b-- > 0;) {
    const thmm_primitive_batch &batch = p.batches[b];
    switch (batch.primitive) {
      case THMM_RAKE: d.ReverseRakes(p.rakes + batch.offset, batch.count); break;
      case THMM_BRANCH_COMBINATION: d.ReverseCombine(p.combinations + batch.offset, batch.count); break;
      case THMM_BRANCH_ABSORPTION: d.ReverseAbsorb(p.absorptions + batch.offset, batch.count); break;
      case THMM_COMPRESSION: d.ReverseCompress(p.compressions + batch.offset, batch.count); break;
    }
  }
}

// Expansion: roots first, then rake/compression batches in reverse
// (src/inference.cc:641-667,345-412).
template <class D> void Expand(const thmm_plan_arrays &p, D &d) {
  for (std::size_t b = p.num_batches; b-- > 0;) {
    const thmm_primitive_batch &batch = p.batches[b];
    if (batch.primitive == THMM_RAKE) d.ExpandRakes(p.rakes + batch.offset, batch.count);
    if (batch.primitive == THMM_COMPRESSION) d.ExpandCompress(p.compressions + batch.offset, batch.count);
  }
}

// src/inference.cc:148-245: linear sum-product, no scaling.
struct SumProduct {
  View v;
  std::vector<double> nodes, paths, branches, scratch;
  SumProduct(const thmm_plan_arrays &p, std::size_t k, const double *n, const double *e)
      : v(p, k), nodes(n, n + p.num_nodes * k), paths(e, e + p.num_edges * k * k),
        branches(p.num_branches * k), scratch(k * k) {}
  void Rake(const thmm_rake *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      const double *path = &paths[ops[i].edge * v.kk];
      const double *leaf = &nodes[ops[i].leaf * v.k];
      double *out = &branches[ops[i].branch * v.k];
      for (std::size_t a = 0; a < v.k; ++a) {
        double s = 0.0;
        for (std::size_t c = 0; c < v.k; ++c) s += path[a * v.k + c] * leaf[c];
        out[a] = s;
      }
    }
  }
  void Combine(const thmm_branch_combination *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i)
      for (std::size_t s = 0; s < v.k; ++s)
        branches[ops[i].destination * v.k + s] *= branches[ops[i].source * v.k + s];
  }
  void Absorb(const thmm_branch_absorption *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i)
      for (std::size_t s = 0; s < v.k; ++s)
        nodes[ops[i].parent * v.k + s] *= branches[ops[i].branch * v.k + s];
  }
  void Compress(const thmm_compression *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      double *left = &paths[ops[i].left_edge * v.kk];
      const double *mid = &nodes[ops[i].middle * v.k];
      const double *right = &paths[ops[i].right_edge * v.kk];
      std::copy(left, left + v.kk, scratch.begin());
      for (std::size_t a = 0; a < v.k; ++a)
        for (std::size_t c = 0; c < v.k; ++c) {
          double s = 0.0;
          for (std::size_t m = 0; m < v.k; ++m) s += scratch[a * v.k + m] * mid[m] * right[m * v.k + c];
          left[a * v.k + c] = s;
        }
    }
  }
};

// src/inference.cc:697-836: linear sum-product with per-op max
// normalisation and additive log scales.
struct Scaled {
  View v;
  std::vector<double> nodes, paths, branches, scratch, node_scale, path_scale, branch_scale;
  Scaled(const thmm_plan_arrays &p, std::size_t k, const double *n, const double *e)
      : v(p, k), nodes(n, n + p.num_nodes * k), paths(e, e + p.num_edges * k * k),
        branches(p.num_branches * k), scratch(k * k), node_scale(p.num_nodes, 0.0),
        path_scale(p.num_edges, 0.0), branch_scale(p.num_branches, 0.0) {}
  // src/inference.cc:787-797
  static double Normalize(double *x, std::size_t size, double input_scale) {
    const double top = *std::max_element(x, x + size);
    if (!(top > 0.0)) return -kInf;
    const double inverse = 1.0 / top;
    for (std::size_t i = 0; i < size; ++i) x[i] *= inverse;
    return input_scale + std::log(top);
  }
  void Rake(const thmm_rake *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      const thmm_rake &op = ops[i];
      const double *path = &paths[op.edge * v.kk];
      const double *leaf = &nodes[op.leaf * v.k];
      double *out = &branches[op.branch * v.k];
      for (std::size_t a = 0; a < v.k; ++a) {
        double s = 0.0;
        for (std::size_t c = 0; c < v.k; ++c) s += path[a * v.k + c] * leaf[c];
        out[a] = s;
      }
      branch_scale[op.branch] = Normalize(out, v.k, path_scale[op.edge] + node_scale[op.leaf]);
    }
  }
  void Combine(const thmm_branch_combination *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      double *dst = &branches[ops[i].destination * v.k];
      const double *src = &branches[ops[i].source * v.k];
      for (std::size_t s = 0; s < v.k; ++s) dst[s] *= src[s];
      branch_scale[ops[i].destination] = Normalize(
          dst, v.k, branch_scale[ops[i].destination] + branch_scale[ops[i].source]);
    }
  }
  void Absorb(const thmm_branch_absorption *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      double *node = &nodes[ops[i].parent * v.k];
      const double *branch = &branches[ops[i].branch * v.k];
      for (std::size_t s = 0; s < v.k; ++s) node[s] *= branch[s];
      node_scale[ops[i].parent] =
          Normalize(node, v.k, node_scale[ops[i].parent] + branch_scale[ops[i].branch]);
    }
  }
  void Compress(const thmm_compression *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      const thmm_compression &op = ops[i];
      double *left = &paths[op.left_edge * v.kk];
      const double *mid = &nodes[op.middle * v.k];
      const double *right = &paths[op.right_edge * v.kk];
      std::copy(left, left + v.kk, scratch.begin());
      for (std::size_t a = 0; a < v.k; ++a)
        for (std::size_t c = 0; c < v.k; ++c) {
          double s = 0.0;
          for (std::size_t m = 0; m < v.k; ++m) s += scratch[a * v.k + m] * mid[m] * right[m * v.k + c];
          left[a * v.k + c] = s;
        }
      path_scale[op.left_edge] = Normalize(
          left, v.kk, path_scale[op.left_edge] + node_scale[op.middle] + path_scale[op.right_edge]);
    }
  }
  // src/inference.cc:778-784
  double Finish(std::uint32_t root) const {
    const double *x = &nodes[root * v.k];
    double sum = 0.0;
    for (std::size_t s = 0; s < v.k; ++s) sum += x[s];
    if (!(sum > 0.0)) return -kInf;
    return node_scale[root] + std::log(sum);
  }
};

// src/inference.cc:247-562: log-domain sum-product with tapes, analytic
// reverse (marginals) and sampling expansion.
struct LogSumProduct {
  View v;
  std::vector<double> nodes, paths, branches;
  std::vector<double> node_adj, path_adj, branch_adj;
  std::vector<double> rake_paths, rake_leaves, comp_left, comp_mid, comp_right, scratch;
  std::vector<double> log_z;
  std::vector<std::uint32_t> assignment;
  const double *uniforms = nullptr;
  double *distance = nullptr;

  LogSumProduct(const thmm_plan_arrays &p, std::size_t k, const double *n, const double *e)
      : v(p, k), nodes(p.num_nodes * k), paths(p.num_edges * k * k), branches(p.num_branches * k),
        node_adj(p.num_nodes * k, 0.0), path_adj(p.num_edges * k * k, 0.0),
        branch_adj(p.num_branches * k, 0.0), rake_paths(p.num_branches * k * k),
        rake_leaves(p.num_branches * k), comp_left(p.num_compressions * k * k),
        comp_mid(p.num_compressions * k), comp_right(p.num_compressions * k * k),
        scratch(k * k), log_z(p.num_roots), assignment(p.num_nodes, 0) {
    // src/inference.cc:249-257: potentials to the log domain.
    for (std::size_t i = 0; i < nodes.size(); ++i) nodes[i] = PotentialLog(n[i]);
    for (std::size_t i = 0; i < paths.size(); ++i) paths[i] = PotentialLog(e[i]);
  }
  // src/inference.cc:259-277
  void Rake(const thmm_rake *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      const thmm_rake &op = ops[i];
      const double *path = &paths[op.edge * v.kk];
      const double *leaf = &nodes[op.leaf * v.k];
      std::copy(path, path + v.kk, &rake_paths[op.branch * v.kk]);
      std::copy(leaf, leaf + v.k, &rake_leaves[op.branch * v.k]);
      for (std::size_t a = 0; a < v.k; ++a)
        branches[op.branch * v.k + a] = Lse(v.k, [&](std::size_t c) { return path[a * v.k + c] + leaf[c]; });
    }
  }
  // src/inference.cc:279-295
  void Combine(const thmm_branch_combination *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i)
      for (std::size_t s = 0; s < v.k; ++s)
        branches[ops[i].destination * v.k + s] += branches[ops[i].source * v.k + s];
  }
  void Absorb(const thmm_branch_absorption *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i)
      for (std::size_t s = 0; s < v.k; ++s)
        nodes[ops[i].parent * v.k + s] += branches[ops[i].branch * v.k + s];
  }
  // src/inference.cc:297-324
  void Compress(const thmm_compression *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      const thmm_compression &op = ops[i];
      double *left = &paths[op.left_edge * v.kk];
      double *sl = &comp_left[op.tape * v.kk];
      double *sm = &comp_mid[op.tape * v.k];
      double *sr = &comp_right[op.tape * v.kk];
      std::copy(left, left + v.kk, sl);
      std::copy(&nodes[op.middle * v.k], &nodes[op.middle * v.k] + v.k, sm);
      std::copy(&paths[op.right_edge * v.kk], &paths[op.right_edge * v.kk] + v.kk, sr);
      for (std::size_t a = 0; a < v.k; ++a)
        for (std::size_t c = 0; c < v.k; ++c)
          left[a * v.k + c] =
              Lse(v.k, [&](std::size_t m) { return sl[a * v.k + m] + sm[m] + sr[m * v.k + c]; });
    }
  }
  // src/inference.cc:326-336 (per root)
  void Finish() {
    for (std::size_t r = 0; r < v.p.num_roots; ++r) {
      const double *root = &nodes[v.p.roots[r] * v.k];
      log_z[r] = Lse(v.k, [&](std::size_t s) { return root[s]; });
      if (!std::isfinite(log_z[r])) throw DomainError("the tree HMM has a nonpositive partition function");
    }
  }
  // src/inference.cc:338-343
  void SeedAdjoint() {
    for (std::size_t r = 0; r < v.p.num_roots; ++r) {
      const std::uint32_t root = v.p.roots[r];
      for (std::size_t s = 0; s < v.k; ++s)
        node_adj[root * v.k + s] = std::exp(nodes[root * v.k + s] - log_z[r]);
    }
  }
  // src/inference.cc:416-464
  void ReverseCompress(const thmm_compression *ops, std::size_t count) {
    for (std::size_t i = count; i-- > 0;) {
      const thmm_compression &op = ops[i];
      const double *l = &comp_left[op.tape * v.kk];
      const double *m = &comp_mid[op.tape * v.k];
      const double *r = &comp_right[op.tape * v.kk];
      double *out_adj = &path_adj[op.left_edge * v.kk];
      double *mid_adj = &node_adj[op.middle * v.k];
      double *right_adj = &path_adj[op.right_edge * v.kk];
      std::fill(scratch.begin(), scratch.end(), 0.0);
      for (std::size_t a = 0; a < v.k; ++a)
        for (std::size_t c = 0; c < v.k; ++c) {
          const double adj = out_adj[a * v.k + c];
          const double out = Lse(v.k, [&](std::size_t x) { return l[a * v.k + x] + m[x] + r[x * v.k + c]; });
          if (adj == 0.0 || !std::isfinite(out)) continue;
          for (std::size_t x = 0; x < v.k; ++x) {
            const double term = l[a * v.k + x] + m[x] + r[x * v.k + c];
            if (!std::isfinite(term)) continue;
            const double contribution = adj * std::exp(term - out);
            scratch[a * v.k + x] += contribution;
            mid_adj[x] += contribution;
            right_adj[x * v.k + c] += contribution;
          }
        }
      std::copy(scratch.begin(), scratch.end(), out_adj);
    }
  }
  // src/inference.cc:466-475
  void ReverseAbsorb(const thmm_branch_absorption *ops, std::size_t count) {
    for (std::size_t i = count; i-- > 0;)
      for (std::size_t s = 0; s < v.k; ++s)
        branch_adj[ops[i].branch * v.k + s] += node_adj[ops[i].parent * v.k + s];
  }
  // src/inference.cc:477-486
  void ReverseCombine(const thmm_branch_combination *ops, std::size_t count) {
    for (std::size_t i = count; i-- > 0;)
      for (std::size_t s = 0; s < v.k; ++s)
        branch_adj[ops[i].source * v.k + s] += branch_adj[ops[i].destination * v.k + s];
  }
  // src/inference.cc:488-518
  void ReverseRakes(const thmm_rake *ops, std::size_t count) {
    for (std::size_t i = count; i-- > 0;) {
      const thmm_rake &op = ops[i];
      const double *path = &rake_paths[op.branch * v.kk];
      const double *leaf = &rake_leaves[op.branch * v.k];
      const double *msg_adj = &branch_adj[op.branch * v.k];
      double *p_adj = &path_adj[op.edge * v.kk];
      double *l_adj = &node_adj[op.leaf * v.k];
      for (std::size_t a = 0; a < v.k; ++a) {
        const double message = Lse(v.k, [&](std::size_t c) { return path[a * v.k + c] + leaf[c]; });
        if (!std::isfinite(message)) continue;
        for (std::size_t c = 0; c < v.k; ++c) {
          const double term = path[a * v.k + c] + leaf[c];
          if (!std::isfinite(term)) continue;
          const double contribution = msg_adj[a] * std::exp(term - message);
          p_adj[a * v.k + c] += contribution;
          l_adj[c] += contribution;
        }
      }
    }
  }
  // src/inference.cc:345-361 (per root)
  void SeedSample() {
    for (std::size_t r = 0; r < v.p.num_roots; ++r) {
      const std::uint32_t root = v.p.roots[r];
      const double *x = &nodes[root * v.k];
      assignment[root] = static_cast<std::uint32_t>(
          SampleLog(v.k, [&](std::size_t s) { return x[s]; }, log_z[r], uniforms[root],
                    distance ? &distance[root] : nullptr));
    }
  }
  // src/inference.cc:363-384
  void ExpandRakes(const thmm_rake *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      const thmm_rake &op = ops[i];
      const std::size_t ps = assignment[op.parent];
      const double *path = &rake_paths[op.branch * v.kk];
      const double *leaf = &rake_leaves[op.branch * v.k];
      for (std::size_t c = 0; c < v.k; ++c) scratch[c] = path[ps * v.k + c] + leaf[c];
      const double norm = Lse(v.k, [&](std::size_t c) { return scratch[c]; });
      assignment[op.leaf] = static_cast<std::uint32_t>(
          SampleLog(v.k, [&](std::size_t c) { return scratch[c]; }, norm, uniforms[op.leaf],
                    distance ? &distance[op.leaf] : nullptr));
    }
  }
  // src/inference.cc:386-412
  void ExpandCompress(const thmm_compression *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      const thmm_compression &op = ops[i];
      const std::size_t ps = assignment[op.parent];
      const std::size_t cs = assignment[op.child];
      const double *l = &comp_left[op.tape * v.kk];
      const double *m = &comp_mid[op.tape * v.k];
      const double *r = &comp_right[op.tape * v.kk];
      for (std::size_t x = 0; x < v.k; ++x) scratch[x] = l[ps * v.k + x] + m[x] + r[x * v.k + cs];
      const double norm = Lse(v.k, [&](std::size_t x) { return scratch[x]; });
      assignment[op.middle] = static_cast<std::uint32_t>(
          SampleLog(v.k, [&](std::size_t x) { return scratch[x]; }, norm, uniforms[op.middle],
                    distance ? &distance[op.middle] : nullptr));
    }
  }
};

// src/inference.cc:564-695: log-domain max-product with choice tapes.
struct MaxProduct {
  View v;
  std::vector<double> nodes, paths, branches, scratch;
  std::vector<std::uint32_t> rake_choice, comp_choice, assignment;
  MaxProduct(const thmm_plan_arrays &p, std::size_t k, const double *n, const double *e)
      : v(p, k), nodes(p.num_nodes * k), paths(p.num_edges * k * k), branches(p.num_branches * k),
        scratch(k * k), rake_choice(p.num_branches * k), comp_choice(p.num_compressions * k * k),
        assignment(p.num_nodes, 0) {
    for (std::size_t i = 0; i < nodes.size(); ++i) nodes[i] = PotentialLog(n[i]);
    for (std::size_t i = 0; i < paths.size(); ++i) paths[i] = PotentialLog(e[i]);
  }
  void Rake(const thmm_rake *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      const thmm_rake &op = ops[i];
      const double *path = &paths[op.edge * v.kk];
      const double *leaf = &nodes[op.leaf * v.k];
      for (std::size_t a = 0; a < v.k; ++a)
        rake_choice[op.branch * v.k + a] = static_cast<std::uint32_t>(ArgMax(
            v.k, [&](std::size_t c) { return path[a * v.k + c] + leaf[c]; },
            &branches[op.branch * v.k + a]));
    }
  }
  void Combine(const thmm_branch_combination *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i)
      for (std::size_t s = 0; s < v.k; ++s)
        branches[ops[i].destination * v.k + s] += branches[ops[i].source * v.k + s];
  }
  void Absorb(const thmm_branch_absorption *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i)
      for (std::size_t s = 0; s < v.k; ++s)
        nodes[ops[i].parent * v.k + s] += branches[ops[i].branch * v.k + s];
  }
  // Candidate association (L + mid) + R as at src/inference.cc:630-633.
  void Compress(const thmm_compression *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i) {
      const thmm_compression &op = ops[i];
      double *left = &paths[op.left_edge * v.kk];
      const double *mid = &nodes[op.middle * v.k];
      const double *right = &paths[op.right_edge * v.kk];
      std::copy(left, left + v.kk, scratch.begin());
      for (std::size_t a = 0; a < v.k; ++a)
        for (std::size_t c = 0; c < v.k; ++c)
          comp_choice[op.tape * v.kk + a * v.k + c] = static_cast<std::uint32_t>(ArgMax(
              v.k, [&](std::size_t m) { return scratch[a * v.k + m] + mid[m] + right[m * v.k + c]; },
              &left[a * v.k + c]));
    }
  }
  void ExpandRakes(const thmm_rake *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i)
      assignment[ops[i].leaf] = rake_choice[ops[i].branch * v.k + assignment[ops[i].parent]];
  }
  void ExpandCompress(const thmm_compression *ops, std::size_t count) {
    for (std::size_t i = 0; i < count; ++i)
      assignment[ops[i].middle] =
          comp_choice[ops[i].tape * v.kk + assignment[ops[i].parent] * v.k + assignment[ops[i].child]];
  }
};

template <class F> int Guard(F body) {
  try {
    body();
    return 0;
  } catch (const DomainError &) {
    return 4;
  } catch (const std::exception &) {
    return 1;
  }
}

std::size_t EnumerationCount(std::size_t states, std::size_t nodes) {
  std::size_t total = 1;
  for (std::size_t i = 0; i < nodes; ++i) {
    if (total > (std::size_t{1} << 26) / states) throw InvalidArgument("brute force is too large");
    total *= states;
  }
  return total;
}

double Weight(const thmm_plan_arrays &p, std::size_t k, const double *nodes, const double *edges,
              const std::uint32_t *x) {
  double w = 1.0;
  for (std::size_t n = 0; n < p.num_nodes; ++n) w *= nodes[n * k + x[n]];
  for (std::size_t e = 0; e < p.num_edges; ++e)
    w *= edges[(e * k + x[p.edge_parents[e]]) * k + x[p.edge_children[e]]];
  return w;
}

void Decode(std::size_t code, std::size_t k, std::size_t n, std::uint32_t *x) {
  for (std::size_t i = 0; i < n; ++i) {
    x[i] = static_cast<std::uint32_t>(code % k);
    code /= k;
  }
}

}  // namespace

extern "C" {

int oracle_partition(const thmm_plan_arrays *plan, size_t states, const double *nodes,
                     const double *edges, double *z) {
  return Guard([&] {
    SumProduct d(*plan, states, nodes, edges);
    Contract(*plan, d);
    // src/inference.cc:216-220 (per root)
    for (std::size_t r = 0; r < plan->num_roots; ++r) {
      double sum = 0.0;
      for (std::size_t s = 0; s < states; ++s) sum += d.nodes[plan->roots[r] * states + s];
      z[r] = sum;
    }
  });
}

int oracle_log_partition(const thmm_plan_arrays *plan, size_t states, const double *nodes,
                         const double *edges, double *log_z) {
  return Guard([&] {
    Scaled d(*plan, states, nodes, edges);
    Contract(*plan, d);
    for (std::size_t r = 0; r < plan->num_roots; ++r) log_z[r] = d.Finish(plan->roots[r]);
  });
}

int oracle_marginals(const thmm_plan_arrays *plan, size_t states, const double *nodes,
                     const double *edges, double *log_z, double *node_marginals,
                     double *edge_marginals) {
  return Guard([&] {
    LogSumProduct d(*plan, states, nodes, edges);
    Contract(*plan, d);
    d.Finish();
    d.SeedAdjoint();
    Reverse(*plan, d);
    std::copy(d.log_z.begin(), d.log_z.end(), log_z);
    std::copy(d.node_adj.begin(), d.node_adj.end(), node_marginals);
    std::copy(d.path_adj.begin(), d.path_adj.end(), edge_marginals);
  });
}

int oracle_map(const thmm_plan_arrays *plan, size_t states, const double *nodes,
               const double *edges, double *log_weight, uint32_t *assignment) {
  return Guard([&] {
    MaxProduct d(*plan, states, nodes, edges);
    Contract(*plan, d);
    // src/inference.cc:641-649 (per root)
    for (std::size_t r = 0; r < plan->num_roots; ++r) {
      const std::uint32_t root = plan->roots[r];
      const double *x = &d.nodes[root * states];
      d.assignment[root] =
          static_cast<std::uint32_t>(ArgMax(states, [&](std::size_t s) { return x[s]; }, &log_weight[r]));
      if (!std::isfinite(log_weight[r])) throw DomainError("the tree HMM has no positive-weight assignment");
    }
    Expand(*plan, d);
    std::copy(d.assignment.begin(), d.assignment.end(), assignment);
  });
}

int oracle_sample(const thmm_plan_arrays *plan, size_t states, const double *nodes,
                  const double *edges, const double *uniforms, uint32_t *assignment,
                  double *boundary_distance) {
  return Guard([&] {
    // src/inference.cc:350-355
    for (std::size_t i = 0; i < plan->num_nodes; ++i)
      if (!std::isfinite(uniforms[i]) || uniforms[i] < 0.0 || uniforms[i] >= 1.0)
        throw InvalidArgument("posterior-sampling variates must lie in [0, 1)");
    LogSumProduct d(*plan, states, nodes, edges);
    Contract(*plan, d);
    d.Finish();
    d.uniforms = uniforms;
    d.distance = boundary_distance;
    if (boundary_distance)
      std::fill(boundary_distance, boundary_distance + plan->num_nodes, kInf);
    d.SeedSample();
    Expand(*plan, d);
    std::copy(d.assignment.begin(), d.assignment.end(), assignment);
  });
}

double oracle_assignment_weight(const thmm_plan_arrays *plan, size_t states, const double *nodes,
                                const double *edges, const uint32_t *assignment) {
  return Weight(*plan, states, nodes, edges, assignment);
}

int oracle_brute_force(const thmm_plan_arrays *plan, size_t states, const double *nodes,
                       const double *edges, double *z, double *node_marginals,
                       double *edge_marginals, double *max_weight, uint32_t *max_assignment) {
  return Guard([&] {
    const thmm_plan_arrays &p = *plan;
    const std::size_t k = states;
    const std::size_t total = EnumerationCount(k, p.num_nodes);
    std::vector<std::uint32_t> x(p.num_nodes), best(p.num_nodes, 0);
    std::vector<double> nm(p.num_nodes * k, 0.0), em(p.num_edges * k * k, 0.0);
    double sum = 0.0, top = 0.0;
    for (std::size_t code = 0; code < total; ++code) {
      Decode(code, k, p.num_nodes, x.data());
      const double w = Weight(p, k, nodes, edges, x.data());
      sum += w;
      for (std::size_t n = 0; n < p.num_nodes; ++n) nm[n * k + x[n]] += w;
      for (std::size_t e = 0; e < p.num_edges; ++e)
        em[(e * k + x[p.edge_parents[e]]) * k + x[p.edge_children[e]]] += w;
      if (w > top) {
        top = w;
        best = x;
      }
    }
    if (z) *z = sum;
    if (node_marginals)
      for (std::size_t i = 0; i < nm.size(); ++i) node_marginals[i] = nm[i] / sum;
    if (edge_marginals)
      for (std::size_t i = 0; i < em.size(); ++i) edge_marginals[i] = em[i] / sum;
    if (max_weight) *max_weight = top;
    if (max_assignment) std::copy(best.begin(), best.end(), max_assignment);
  });
}

int oracle_brute_conditional_sample(const thmm_plan_arrays *plan, size_t states,
                                    const double *nodes, const double *edges,
                                    const double *uniforms, uint32_t *assignment) {
  return Guard([&] {
    const thmm_plan_arrays &p = *plan;
    const std::size_t k = states;
    const std::size_t total = EnumerationCount(k, p.num_nodes);
    std::vector<std::uint32_t> order;
    for (std::size_t r = 0; r < p.num_roots; ++r) order.push_back(p.roots[r]);
    for (std::size_t b = p.num_batches; b-- > 0;) {
      const thmm_primitive_batch &batch = p.batches[b];
      for (std::size_t i = batch.offset; i < batch.offset + batch.count; ++i) {
        if (batch.primitive == THMM_RAKE) order.push_back(p.rakes[i].leaf);
        if (batch.primitive == THMM_COMPRESSION) order.push_back(p.compressions[i].middle);
      }
    }
    std::vector<int> fixed(p.num_nodes, -1);
    std::vector<std::uint32_t> x(p.num_nodes);
    std::vector<double> mass(k);
    for (std::uint32_t node : order) {
      std::fill(mass.begin(), mass.end(), 0.0);
      for (std::size_t code = 0; code < total; ++code) {
        Decode(code, k, p.num_nodes, x.data());
        bool consistent = true;
        for (std::size_t n = 0; n < p.num_nodes && consistent; ++n)
          consistent = fixed[n] < 0 || static_cast<std::uint32_t>(fixed[n]) == x[n];
        if (!consistent) continue;
        mass[x[node]] += Weight(p, k, nodes, edges, x.data());
      }
      double sum = 0.0;
      for (double m : mass) sum += m;
      // Inverse CDF in log form, the rule of src/inference.cc:128-146.
      const double log_sum = std::log(sum);
      const std::size_t s = SampleLog(
          k, [&](std::size_t i) { return PotentialLog(mass[i]); }, log_sum, uniforms[node], nullptr);
      fixed[node] = static_cast<int>(s);
    }
    for (std::size_t n = 0; n < p.num_nodes; ++n) assignment[n] = static_cast<std::uint32_t>(fixed[n]);
  });
}

int oracle_marginals_batched(const thmm_plan_arrays *plan, size_t states, size_t items,
                             const double *nodes, const double *edges, double *log_z,
                             double *node_marginals, double *edge_marginals, size_t threads) {
  if (threads == 0) threads = 1;
  const std::size_t nk = plan->num_nodes * states;
  const std::size_t ekk = plan->num_edges * states * states;
  std::vector<int> status(threads, 0);
  std::vector<std::thread> pool;
  for (std::size_t t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      for (std::size_t item = t; item < items; item += threads) {
        const int rc = oracle_marginals(plan, states, nodes + item * nk, edges,
                                        log_z + item * plan->num_roots, node_marginals + item * nk,
                                        edge_marginals + item * ekk);
        if (rc != 0) status[t] = rc;
      }
    });
  }
  for (std::thread &worker : pool) worker.join();
  for (int rc : status)
    if (rc != 0) return rc;
  return 0;
}

}  // extern "C"
