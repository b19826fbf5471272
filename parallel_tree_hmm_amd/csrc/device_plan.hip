// Rake-compress planner on the device. It builds the plan MakePlan builds
// on the host (plan.cc, Contract): the same ops, in the same order, in the
// same batches and rounds, so every kernel later computes the same values.
//
// Per round, over the live non-root nodes (an ascending list):
//  * rakes: the live leaves, compacted in ascending order; branch ids in
//    that order. Child counts and child XORs of the parents are updated
//    with atomics (order-free: a count and an XOR).
//  * combinations: the round's (parent, branch) pairs stably sorted by
//    parent (radix sort), so each parent's branches stay in branch order.
//    Stage t of a parent's pairwise reduction pairs the branches at
//    positions 2j 2^t and (2j + 1) 2^t of its sorted list (the kept
//    branches after t stages are those at multiples of 2^t), so each
//    stage's ops are laid out from a scan of per-parent pair counts.
//  * absorptions: one per parent, its first branch.
//  * compressions: unary live nodes (one live child) form chains; a node's
//    chain head and distance from it come from pointer jumping over
//    "parent is unary" links. The nodes at even distance are spliced out,
//    listed by chain head (ascending) then distance, as the host walk lists
//    them: a scan of per-head counts places each one.
// Cycles are found first, by pointer jumping to the roots (the host's
// "parent links contain a cycle" error).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "plan.h"

namespace thmm {
namespace {

constexpr Index kNone = 0xFFFFFFFFu;
constexpr unsigned kBlock = 256;
constexpr unsigned kMaxBlocks = 1u << 20;

void Check(hipError_t status, const char *what) {
  if (status != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
}

// Grid of a grid-stride loop over `n` items (at most 2^20 blocks, so the
// work-item count of a dispatch stays below 2^32).
unsigned Grid(std::size_t n) {
  const std::size_t blocks = (n + kBlock - 1) / kBlock;
  return static_cast<unsigned>(std::max<std::size_t>(1, std::min<std::size_t>(blocks, kMaxBlocks)));
}

__device__ inline std::size_t Thread() { return static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; }
__device__ inline std::size_t Stride() { return static_cast<std::size_t>(gridDim.x) * blockDim.x; }

// Device array owned by the planner.
template <typename T> class Buffer {
 public:
  Buffer(std::size_t n, const char *what) { Check(hipMalloc(&p_, std::max<std::size_t>(n, 1) * sizeof(T)), what); }
  ~Buffer() {
    if (p_ != nullptr) static_cast<void>(hipFree(p_));
  }
  Buffer(const Buffer &) = delete;
  Buffer &operator=(const Buffer &) = delete;
  T *get() const { return p_; }
  void swap(Buffer &o) { std::swap(p_, o.p_); }

 private:
  T *p_ = nullptr;
};

// ---- kernels ----

// par[i] = parent or kNone; nonroot[i]; bad = the smallest node whose parent
// is out of range (or itself).
__global__ void ReadParents(const std::int64_t *parents, std::size_t n, Index *par, Index *nonroot, Index *bad) {
  for (std::size_t i = Thread(); i < n; i += Stride()) {
    const std::int64_t p = parents[i];
    if (p == -1) {
      par[i] = kNone;
      nonroot[i] = 0;
    } else if (p < 0 || static_cast<std::size_t>(p) >= n || static_cast<std::size_t>(p) == i) {
      par[i] = kNone;
      nonroot[i] = 0;
      atomicMin(bad, static_cast<Index>(i));
    } else {
      par[i] = static_cast<Index>(p);
      nonroot[i] = 1;
    }
  }
}

// Root finding by pointer jumping: r = the parent (a root: itself), then
// r = r[r] until nothing changes or 32 steps have passed (any acyclic depth
// below 2^32 converges). A node whose r is not a root then lies on or above
// a cycle.
__global__ void RootInit(const Index *par, std::size_t n, Index *r) {
  for (std::size_t i = Thread(); i < n; i += Stride()) r[i] = par[i] == kNone ? static_cast<Index>(i) : par[i];
}

__global__ void RootJump(const Index *r, std::size_t n, Index *r_next, Index *changed) {
  for (std::size_t i = Thread(); i < n; i += Stride()) {
    const Index a = r[i], b = r[a];
    r_next[i] = b;
    if (b != a) *changed = 1;
  }
}

__global__ void CycleCheck(const Index *par, const Index *r, std::size_t n, Index *cyclic) {
  for (std::size_t i = Thread(); i < n; i += Stride())
    if (par[r[i]] != kNone) *cyclic = 1;
}

// Canonical edges (one per non-root node, ascending), slots, child counts and
// child XORs.
__global__ void Edges(const Index *par, const Index *edge_index, std::size_t n, Index *edge_parents,
                      Index *edge_children, Index *slot, Index *cnt, Index *cxor, std::uint8_t *alive) {
  for (std::size_t i = Thread(); i < n; i += Stride()) {
    const Index p = par[i];
    alive[i] = 1;
    if (p == kNone) {
      slot[i] = kNone;
      continue;
    }
    const Index e = edge_index[i];
    edge_parents[e] = p;
    edge_children[e] = static_cast<Index>(i);
    slot[i] = e;
    atomicAdd(&cnt[p], 1u);
    atomicXor(&cxor[p], static_cast<Index>(i));
  }
}

// One rake per live leaf; the parents lose the leaf.
__global__ void Rakes(const Index *leaves, Index count, const Index *par, const Index *slot, Index first_branch,
                      thmm_rake *out, Index *keys, Index *vals, Index *cnt, Index *cxor, std::uint8_t *alive) {
  for (std::size_t r = Thread(); r < count; r += Stride()) {
    const Index leaf = leaves[r], p = par[leaf];
    const Index branch = first_branch + static_cast<Index>(r);
    out[r] = {leaf, p, slot[leaf], branch};
    keys[r] = p;
    vals[r] = branch;
    atomicSub(&cnt[p], 1u);
    atomicXor(&cxor[p], leaf);
    alive[leaf] = 0;
  }
}

// Group sizes from group starts; absorptions (parent, first branch).
__global__ void Groups(const Index *start, Index groups, Index total, const Index *skeys, const Index *svals,
                       Index round, Index *size, thmm_branch_absorption *absorb) {
  for (std::size_t g = Thread(); g < groups; g += Stride()) {
    const Index s = start[g];
    size[g] = (g + 1 < groups ? start[g + 1] : total) - s;
    absorb[g] = {skeys[s], svals[s], round};
  }
}

// Pairs of a group of `size` branches at stage t: floor(ceil(size / 2^t) / 2).
__device__ inline Index StagePairs(Index size, unsigned t) {
  const Index left = static_cast<Index>((static_cast<std::uint64_t>(size) + (1ull << t) - 1) >> t);
  return left / 2;
}

__global__ void StageCounts(const Index *size, Index groups, unsigned t, Index *pairs) {
  for (std::size_t g = Thread(); g < groups; g += Stride()) pairs[g] = StagePairs(size[g], t);
}

// Combination q of stage t: its group is the first whose inclusive pair count
// exceeds q.
__global__ void StageCombinations(const Index *inclusive, const Index *size, const Index *start, Index groups,
                                  Index total, unsigned t, const Index *svals, thmm_branch_combination *out) {
  for (std::size_t q = Thread(); q < total; q += Stride()) {
    Index lo = 0, hi = groups;
    while (lo < hi) {
      const Index mid = lo + (hi - lo) / 2;
      if (inclusive[mid] > q) hi = mid; else lo = mid + 1;
    }
    const Index g = lo;
    const Index j = static_cast<Index>(q) - (inclusive[g] - StagePairs(size[g], t));
    const std::uint64_t base = start[g];
    out[q] = {svals[base + ((2ull * j) << t)], svals[base + ((2ull * j + 1) << t)], static_cast<Index>(t)};
  }
}

__global__ void MarkUnary(const Index *live, Index count, const std::uint8_t *alive, const Index *cnt,
                          std::uint8_t *unary) {
  for (std::size_t k = Thread(); k < count; k += Stride()) {
    const Index i = live[k];
    unary[i] = alive[i] != 0 && cnt[i] == 1 ? 1 : 0;
  }
}

__global__ void ClearUnary(const Index *nodes, Index count, std::uint8_t *unary) {
  for (std::size_t k = Thread(); k < count; k += Stride()) unary[nodes[k]] = 0;
}

// up = the parent when it is unary (the node is inside a chain), else the
// node itself (a chain head); d = 1 or 0.
__global__ void ChainInit(const Index *nodes, Index count, const Index *par, const std::uint8_t *unary, Index *up,
                          Index *d) {
  for (std::size_t k = Thread(); k < count; k += Stride()) {
    const Index i = nodes[k], p = par[i];
    const bool inner = p != kNone && unary[p] != 0;
    up[i] = inner ? p : i;
    d[i] = inner ? 1 : 0;
  }
}

// One pointer-jumping step: distances add up along the jump.
__global__ void ChainJump(const Index *nodes, Index count, const Index *up, const Index *d, Index *up_next,
                          Index *d_next, Index *changed) {
  for (std::size_t k = Thread(); k < count; k += Stride()) {
    const Index i = nodes[k], a = up[i];
    if (up[a] != a) {
      up_next[i] = up[a];
      d_next[i] = d[i] + d[a];
      *changed = 1;
    } else {
      up_next[i] = a;
      d_next[i] = d[i];
    }
  }
}

__global__ void HeadIndex(const Index *heads, Index count, Index *head_index, Index *length) {
  for (std::size_t h = Thread(); h < count; h += Stride()) {
    head_index[heads[h]] = static_cast<Index>(h);
    length[h] = 0;
  }
}

__global__ void ChainLength(const Index *nodes, Index count, const Index *up, const Index *d,
                            const Index *head_index, Index *length) {
  for (std::size_t k = Thread(); k < count; k += Stride()) {
    const Index i = nodes[k];
    atomicMax(&length[head_index[up[i]]], d[i] + 1);
  }
}

__global__ void SelectedPerHead(const Index *length, Index count, Index *selected) {
  for (std::size_t h = Thread(); h < count; h += Stride()) selected[h] = (length[h] + 1) / 2;
}

// The nodes at even distance from their chain head, placed by head then
// distance.
__global__ void Compressions(const Index *nodes, Index count, const Index *up, const Index *d,
                             const Index *head_index, const Index *offset, const Index *cxor, const Index *slot,
                             const Index *par, Index first, thmm_compression *out) {
  for (std::size_t k = Thread(); k < count; k += Stride()) {
    const Index i = nodes[k];
    if (d[i] % 2 != 0) continue;
    const Index s = offset[head_index[up[i]]] + d[i] / 2;
    const Index child = cxor[i];
    out[s] = {slot[i], i, slot[child], par[i], child, first + s};
  }
}

// Splices the compressed nodes out (after every record was read).
__global__ void SpliceCompressions(const thmm_compression *ops, Index count, Index *cxor, Index *par, Index *slot,
                                   std::uint8_t *alive) {
  for (std::size_t s = Thread(); s < count; s += Stride()) {
    const thmm_compression op = ops[s];
    atomicXor(&cxor[op.parent], op.middle ^ op.child);
    par[op.child] = op.parent;
    slot[op.child] = op.left_edge;
    alive[op.middle] = 0;
  }
}

// ---- selection predicates ----

struct IsZero {
  const Index *values;
  __host__ __device__ bool operator()(Index i) const { return values[i] == 0; }
};
struct IsSet {
  const std::uint8_t *flags;
  __host__ __device__ bool operator()(Index i) const { return flags[i] != 0; }
};
struct IsRoot {
  const Index *par;
  __host__ __device__ bool operator()(Index i) const { return par[i] == kNone; }
};
struct IsNonRoot {
  const Index *par;
  __host__ __device__ bool operator()(Index i) const { return par[i] != kNone; }
};
struct StartsGroup {
  const Index *keys;
  __host__ __device__ bool operator()(Index r) const { return r == 0 || keys[r] != keys[r - 1]; }
};
struct IsHead {
  const Index *up;
  __host__ __device__ bool operator()(Index i) const { return up[i] == i; }
};

// hipCUB calls with one shared scratch buffer, grown on demand.
class Primitives {
 public:
  explicit Primitives(hipStream_t s) : stream_(s) {
    Check(hipHostMalloc(reinterpret_cast<void **>(&host_word_), sizeof(Index), hipHostMallocDefault),
          "device planner host word");
  }
  ~Primitives() {
    if (temp_ != nullptr) static_cast<void>(hipFree(temp_));
    if (host_word_ != nullptr) static_cast<void>(hipHostFree(host_word_));
  }

  // out[0..selected) = the indices in [0, n) of `in` (or the indices
  // themselves when `in` is null) whose predicate holds, in order.
  template <typename Pred> Index Select(const Index *in, Index *out, std::size_t n, Pred pred) {
    if (n == 0) return 0;
    std::size_t bytes = 0;
    if (in == nullptr) {
      hipcub::CountingInputIterator<Index> first(0);
      Check(hipcub::DeviceSelect::If(nullptr, bytes, first, out, count_, static_cast<int>(n), pred, stream_),
            "device planner selection");
      Reserve(bytes);
      Check(hipcub::DeviceSelect::If(temp_, bytes, first, out, count_, static_cast<int>(n), pred, stream_),
            "device planner selection");
    } else {
      // The predicate is evaluated on the values of `in`.
      Check(hipcub::DeviceSelect::If(nullptr, bytes, in, out, count_, static_cast<int>(n), pred, stream_),
            "device planner selection");
      Reserve(bytes);
      Check(hipcub::DeviceSelect::If(temp_, bytes, in, out, count_, static_cast<int>(n), pred, stream_),
            "device planner selection");
    }
    return Read(count_, "device planner selection count");
  }

  void ExclusiveSum(const Index *in, Index *out, std::size_t n) {
    if (n == 0) return;
    std::size_t bytes = 0;
    Check(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, static_cast<int>(n), stream_), "device planner scan");
    Reserve(bytes);
    Check(hipcub::DeviceScan::ExclusiveSum(temp_, bytes, in, out, static_cast<int>(n), stream_), "device planner scan");
  }

  void InclusiveSum(const Index *in, Index *out, std::size_t n) {
    if (n == 0) return;
    std::size_t bytes = 0;
    Check(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, static_cast<int>(n), stream_), "device planner scan");
    Reserve(bytes);
    Check(hipcub::DeviceScan::InclusiveSum(temp_, bytes, in, out, static_cast<int>(n), stream_), "device planner scan");
  }

  Index Max(const Index *in, std::size_t n) {
    if (n == 0) return 0;
    std::size_t bytes = 0;
    Check(hipcub::DeviceReduce::Max(nullptr, bytes, in, count_, static_cast<int>(n), stream_), "device planner reduction");
    Reserve(bytes);
    Check(hipcub::DeviceReduce::Max(temp_, bytes, in, count_, static_cast<int>(n), stream_), "device planner reduction");
    return Read(count_, "device planner reduction result");
  }

  // Stable sort of (key, value) pairs by the low `bits` bits of the keys.
  void SortPairs(const Index *keys, Index *keys_out, const Index *vals, Index *vals_out, std::size_t n, int bits) {
    if (n == 0) return;
    std::size_t bytes = 0;
    Check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, keys_out, vals, vals_out, static_cast<int>(n), 0,
                                             bits, stream_),
          "device planner sort");
    Reserve(bytes);
    Check(hipcub::DeviceRadixSort::SortPairs(temp_, bytes, keys, keys_out, vals, vals_out, static_cast<int>(n), 0,
                                             bits, stream_),
          "device planner sort");
  }

  // One device word, read after the stream's work so far (through a pinned
  // host word).
  Index Read(const Index *word, const char *what = "device planner read") {
    Check(hipMemcpyAsync(host_word_, word, sizeof(Index), hipMemcpyDeviceToHost, stream_), what);
    Check(hipStreamSynchronize(stream_), what);
    return *host_word_;
  }

 private:
  void Reserve(std::size_t bytes) {
    if (bytes <= temp_bytes_) return;
    if (temp_ != nullptr) Check(hipFree(temp_), "device planner scratch");
    temp_ = nullptr;
    temp_bytes_ = 0;
    Check(hipMalloc(&temp_, bytes), "device planner scratch");
    temp_bytes_ = bytes;
  }

  hipStream_t stream_;
  Index *host_word_ = nullptr;
  void *temp_ = nullptr;
  std::size_t temp_bytes_ = 0;
  Buffer<Index> count_buffer_{1, "device planner counter"};
  Index *count_ = count_buffer_.get();
};

template <typename T> std::vector<T> Download(const T *src, std::size_t n, hipStream_t s) {
  std::vector<T> out(n);
  if (n != 0) Check(hipMemcpyAsync(out.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, s), "device plan download");
  return out;
}

// Appends a batch of `count` ops of one kind (skipping empty ones).
void PushBatch(Plan &plan, thmm_primitive primitive, std::size_t offset, std::size_t count) {
  if (count == 0) return;
  plan.batches.push_back({static_cast<std::uint32_t>(primitive), static_cast<std::uint32_t>(offset),
                          static_cast<std::uint32_t>(count)});
}

Plan Contract(const std::int64_t *parents, std::size_t n, hipStream_t s) {
  Plan plan;
  plan.num_nodes = n;
  Primitives prim(s);
  const unsigned G = Grid(n);

  Buffer<Index> par(n, "device planner parents"), nonroot(n, "device planner flags"), bad(1, "device planner flag");
  Check(hipMemsetAsync(bad.get(), 0xFF, sizeof(Index), s), "device planner flag");  // kNone
  ReadParents<<<G, kBlock, 0, s>>>(parents, n, par.get(), nonroot.get(), bad.get());
  Check(hipGetLastError(), "device planner launch");
  const Index first_bad = prim.Read(bad.get(), "device planner range check");
  if (first_bad != kNone) throw std::invalid_argument("parent of node " + std::to_string(first_bad) + " is out of range");

  Buffer<Index> roots(n, "device planner roots");
  const Index num_roots = prim.Select(nullptr, roots.get(), n, IsRoot{par.get()});
  if (num_roots == 0) throw std::invalid_argument("a plan needs at least one root (parent -1)");
  // Cycles are rejected before planning, so every chain of unary nodes met
  // below has a head.
  Buffer<Index> up(n, "device planner chains"), up_next(n, "device planner chains"), changed(1, "device planner flag");
  RootInit<<<G, kBlock, 0, s>>>(par.get(), n, up.get());
  Check(hipGetLastError(), "device planner launch");
  for (int step = 0; step < 32; ++step) {
    Check(hipMemsetAsync(changed.get(), 0, sizeof(Index), s), "device planner clear");
    RootJump<<<G, kBlock, 0, s>>>(up.get(), n, up_next.get(), changed.get());
    Check(hipGetLastError(), "device planner launch");
    up.swap(up_next);
    if (prim.Read(changed.get(), "device planner root jump") == 0) break;
  }
  Check(hipMemsetAsync(changed.get(), 0, sizeof(Index), s), "device planner clear");
  CycleCheck<<<G, kBlock, 0, s>>>(par.get(), up.get(), n, changed.get());
  Check(hipGetLastError(), "device planner launch");
  if (prim.Read(changed.get(), "device planner cycle check") != 0)
    throw std::invalid_argument("parent links contain a cycle");

  // Edges and per-node state.
  Buffer<Index> edge_index(n, "device planner edges"), edge_parents(n, "device planner edges"),
      edge_children(n, "device planner edges"), slot(n, "device planner slots"), cnt(n, "device planner counts"),
      cxor(n, "device planner child sums");
  Buffer<std::uint8_t> alive(n, "device planner flags"), unary(n, "device planner flags");
  prim.ExclusiveSum(nonroot.get(), edge_index.get(), n);
  Check(hipMemsetAsync(cnt.get(), 0, n * sizeof(Index), s), "device planner clear");
  Check(hipMemsetAsync(cxor.get(), 0, n * sizeof(Index), s), "device planner clear");
  Check(hipMemsetAsync(unary.get(), 0, n, s), "device planner clear");
  Edges<<<G, kBlock, 0, s>>>(par.get(), edge_index.get(), n, edge_parents.get(), edge_children.get(), slot.get(),
                             cnt.get(), cxor.get(), alive.get());
  Check(hipGetLastError(), "device planner launch");
  const std::size_t E = n - num_roots;
  plan.num_edges = E;

  // Live non-root nodes (ascending) and op storage: every op kind has at
  // most one op per non-root node.
  Buffer<Index> live(n, "device planner live nodes"), live_next(n, "device planner live nodes");
  Index L = prim.Select(nullptr, live.get(), n, IsNonRoot{par.get()});
  const std::size_t cap = std::max<std::size_t>(E, 1);
  Buffer<thmm_rake> rakes(cap, "device plan rakes");
  Buffer<thmm_branch_combination> combinations(cap, "device plan combinations");
  Buffer<thmm_branch_absorption> absorptions(cap, "device plan absorptions");
  Buffer<thmm_compression> compressions(cap, "device plan compressions");
  Buffer<Index> leaves(cap, "device planner leaves"), keys(cap, "device planner keys"), vals(cap, "device planner keys"),
      skeys(cap, "device planner keys"), svals(cap, "device planner keys"), starts(cap, "device planner groups"),
      sizes(cap, "device planner groups"), pairs(cap, "device planner groups"), inclusive(cap, "device planner groups"),
      chain(cap, "device planner chains"), heads(cap, "device planner chains"), dist(n, "device planner chains"),
      dist_next(n, "device planner chains"), head_index(n, "device planner chains"), length(cap, "device planner chains"),
      per_head(cap, "device planner chains"), head_offset(cap, "device planner chains");
  int key_bits = 1;
  while (key_bits < 32 && (static_cast<std::uint64_t>(n - 1) >> key_bits) != 0) ++key_bits;

  std::size_t num_rakes = 0, num_combinations = 0, num_absorptions = 0, num_compressions = 0;
  Index round = 0;
  while (L != 0) {
    plan.round_batch_offsets.push_back(static_cast<Index>(plan.batches.size()));
    // Rakes.
    const Index R = prim.Select(live.get(), leaves.get(), L, IsZero{cnt.get()});
    if (R == 0) throw std::logic_error("device planner: a round without leaves in an acyclic forest");
    Rakes<<<Grid(R), kBlock, 0, s>>>(leaves.get(), R, par.get(), slot.get(), static_cast<Index>(num_rakes),
                                     rakes.get() + num_rakes, keys.get(), vals.get(), cnt.get(), cxor.get(),
                                     alive.get());
    Check(hipGetLastError(), "device planner launch");
    PushBatch(plan, THMM_RAKE, num_rakes, R);
    num_rakes += R;

    // Branches grouped by parent.
    prim.SortPairs(keys.get(), skeys.get(), vals.get(), svals.get(), R, key_bits);
    const Index groups = prim.Select(nullptr, starts.get(), R, StartsGroup{skeys.get()});
    Groups<<<Grid(groups), kBlock, 0, s>>>(starts.get(), groups, R, skeys.get(), svals.get(), round, sizes.get(),
                                           absorptions.get() + num_absorptions);
    Check(hipGetLastError(), "device planner launch");
    const Index largest = prim.Max(sizes.get(), groups);
    for (unsigned t = 0; t < 32 && (static_cast<std::uint64_t>(1) << t) < largest; ++t) {
      StageCounts<<<Grid(groups), kBlock, 0, s>>>(sizes.get(), groups, t, pairs.get());
      Check(hipGetLastError(), "device planner launch");
      prim.InclusiveSum(pairs.get(), inclusive.get(), groups);
      const Index total = prim.Read(inclusive.get() + groups - 1, "device planner stage count");
      if (total == 0) break;
      StageCombinations<<<Grid(total), kBlock, 0, s>>>(inclusive.get(), sizes.get(), starts.get(), groups, total, t,
                                                       svals.get(), combinations.get() + num_combinations);
      Check(hipGetLastError(), "device planner launch");
      PushBatch(plan, THMM_BRANCH_COMBINATION, num_combinations, total);
      num_combinations += total;
    }
    PushBatch(plan, THMM_BRANCH_ABSORPTION, num_absorptions, groups);
    num_absorptions += groups;

    // Compressions.
    MarkUnary<<<Grid(L), kBlock, 0, s>>>(live.get(), L, alive.get(), cnt.get(), unary.get());
    Check(hipGetLastError(), "device planner launch");
    const Index U = prim.Select(live.get(), chain.get(), L, IsSet{unary.get()});
    Index S = 0;
    if (U != 0) {
      ChainInit<<<Grid(U), kBlock, 0, s>>>(chain.get(), U, par.get(), unary.get(), up.get(), dist.get());
      Check(hipGetLastError(), "device planner launch");
      const Index H = prim.Select(chain.get(), heads.get(), U, IsHead{up.get()});
      for (int step = 0;; ++step) {
        if (step == 33) throw std::logic_error("device planner: chain ranking did not converge");
        Check(hipMemsetAsync(changed.get(), 0, sizeof(Index), s), "device planner clear");
        ChainJump<<<Grid(U), kBlock, 0, s>>>(chain.get(), U, up.get(), dist.get(), up_next.get(), dist_next.get(),
                                             changed.get());
        Check(hipGetLastError(), "device planner launch");
        up.swap(up_next);
        dist.swap(dist_next);
        if (prim.Read(changed.get(), "device planner chain jump") == 0) break;
      }
      HeadIndex<<<Grid(H), kBlock, 0, s>>>(heads.get(), H, head_index.get(), length.get());
      ChainLength<<<Grid(U), kBlock, 0, s>>>(chain.get(), U, up.get(), dist.get(), head_index.get(), length.get());
      SelectedPerHead<<<Grid(H), kBlock, 0, s>>>(length.get(), H, per_head.get());
      Check(hipGetLastError(), "device planner launch");
      prim.ExclusiveSum(per_head.get(), head_offset.get(), H);
      S = prim.Read(head_offset.get() + H - 1, "device planner chain count");
      S += prim.Read(per_head.get() + H - 1, "device planner chain count");
      Compressions<<<Grid(U), kBlock, 0, s>>>(chain.get(), U, up.get(), dist.get(), head_index.get(),
                                              head_offset.get(), cxor.get(), slot.get(), par.get(),
                                              static_cast<Index>(num_compressions),
                                              compressions.get() + num_compressions);
      Check(hipGetLastError(), "device planner launch");
      SpliceCompressions<<<Grid(S), kBlock, 0, s>>>(compressions.get() + num_compressions, S, cxor.get(), par.get(),
                                                    slot.get(), alive.get());
      ClearUnary<<<Grid(U), kBlock, 0, s>>>(chain.get(), U, unary.get());
      Check(hipGetLastError(), "device planner launch");
    }
    PushBatch(plan, THMM_COMPRESSION, num_compressions, S);
    num_compressions += S;

    L = prim.Select(live.get(), live_next.get(), L, IsSet{alive.get()});
    live.swap(live_next);
    ++round;
  }
  plan.round_batch_offsets.push_back(static_cast<Index>(plan.batches.size()));
  plan.num_branches = num_rakes;

  plan.roots = Download(roots.get(), num_roots, s);
  plan.edge_parents = Download(edge_parents.get(), E, s);
  plan.edge_children = Download(edge_children.get(), E, s);
  plan.rakes = Download(rakes.get(), num_rakes, s);
  plan.combinations = Download(combinations.get(), num_combinations, s);
  plan.absorptions = Download(absorptions.get(), num_absorptions, s);
  plan.compressions = Download(compressions.get(), num_compressions, s);
  Check(hipStreamSynchronize(s), "device plan download");
  return plan;
}

}  // namespace

Plan MakePlanOnDevice(const std::int64_t *parents, std::size_t num_nodes, bool parents_on_device, int device) {
  if (num_nodes == 0) throw std::invalid_argument("a plan needs at least one node");
  if (num_nodes >= kNone) throw std::length_error("plan node count exceeds uint32_t");
  // The library scans and sorts count items in int.
  if (num_nodes > static_cast<std::size_t>(std::numeric_limits<int>::max()))
    throw std::length_error("the device planner takes at most 2^31 - 1 nodes");
  Check(hipSetDevice(device), "GPU runtime SetDevice");
  hipStream_t s = nullptr;
  Check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "GPU runtime StreamCreateWithFlags");
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { static_cast<void>(hipStreamDestroy(s)); }
  } guard{s};
  const std::int64_t *device_parents = parents;
  std::unique_ptr<Buffer<std::int64_t>> staged;
  if (!parents_on_device) {
    staged = std::make_unique<Buffer<std::int64_t>>(num_nodes, "device planner parents");
    Check(hipMemcpyAsync(staged->get(), parents, num_nodes * sizeof(std::int64_t), hipMemcpyHostToDevice, s),
          "device planner upload");
    device_parents = staged->get();
  }
  return Contract(device_parents, num_nodes, s);
}

}  // namespace thmm
