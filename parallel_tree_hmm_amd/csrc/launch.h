// Interface between the host executor (workspace.cc) and the gfx950 kernels
// (kernels.hip). Plain structs of device pointers and sizes; no HIP types.
#ifndef TREE_HMM_AMD_LAUNCH_H_
#define TREE_HMM_AMD_LAUNCH_H_

#include <cstddef>
#include <cstdint>

namespace thmm {

// Ops after slot renaming. Slots 0..E-1 hold the (batch-shared) edge
// potentials; compression c writes the fresh per-item slot E + c instead of
// overwriting its left edge, so nothing the reverse pass needs is ever
// overwritten and every slot, node and branch is consumed exactly once.
struct DeviceRake {
  std::uint32_t leaf, slot, branch, parent;
};
struct DeviceCombine {
  std::uint32_t destination, source;
};
struct DeviceAbsorb {
  std::uint32_t parent, branch;
};
struct DeviceCompress {
  std::uint32_t left, middle, right, out;  // out == E + compression index
  std::uint32_t parent, child, pad0, pad1;
};

// Grouped compressions: a connected piece of one compression tree (a root
// compression and the compressions below it down to a fixed height) taken
// as one product S = X_0 D_1 X_1 ... D_{n-1} X_{n-1} over its n input slots
// in tree order; middles[i] is the node between slots[i] and slots[i+1].
// Only `out`, the root compression's slot, is written, so the slots of the
// inner compressions are never stored or read back.
constexpr std::uint32_t kMaxChain = 8;
struct DeviceChain {
  std::uint32_t count, out;
  std::uint32_t slots[kMaxChain];
  std::uint32_t middles[kMaxChain - 1];
  std::uint32_t pad;
};
// Level primitives of grouped compressions (after the plan's 0..3 and the
// fused forward's 4 and 5): 6 for groups of any size, and two reverse
// launches with fewer registers: 7 for a range of groups of 3 or 4 slots,
// 8 for a range of groups of 2 slots.
constexpr std::uint32_t kChainPrimitive = 6;
constexpr std::uint32_t kSmallChainPrimitive = 7;
constexpr std::uint32_t kPairChainPrimitive = 8;

// Max-product compression pieces: a root compression and those of its two
// input slots that are compressions absorbed into it (flag bit 0: the left
// input, the compression left - E; bit 1: the right). An absorbed
// compression's path lives in registers only; its choices are stored as
// usual, so the MAP backtrack is unchanged. Level primitive 9 runs a range
// of pieces.
struct DeviceMaxPiece {
  std::uint32_t root, flags;
};
constexpr std::uint32_t kMaxPiecePrimitive = 9;

// A dispatch's grid size is a 32-bit count of work-items. Kernels that step
// through their arrays by the grid's size take at most 2^20 blocks;
// level launches split their op range so that each dispatch stays below
// 2^31 work-items (kMaxLaunchItems) at up to 64 lanes per (op, item).
inline unsigned GridStrideBlocks(std::uint64_t threads, unsigned block) {
  const std::uint64_t blocks = (threads + block - 1) / block;
  return static_cast<unsigned>(blocks < (1u << 20) ? blocks : (1u << 20));
}
constexpr std::uint64_t kMaxLaunchItems = std::uint64_t{1} << 31;

// Runs launch(offset, count) over consecutive sub-ranges of the level's ops
// with at most kMaxLaunchItems / (batch * 64) ops each (the ops of a level
// are independent and the kernels address them as offset + index).
template <typename F>
inline void InOpChunks(std::uint32_t offset, std::uint32_t count, std::uint64_t batch, F &&launch) {
  const std::uint64_t per = kMaxLaunchItems / (batch * 64);
  const std::uint32_t step = static_cast<std::uint32_t>(per == 0 ? 1 : (per < count ? per : count));
  for (std::uint32_t done = 0; done < count; done += step)
    launch(offset + done, count - done < step ? count - done : step);
}

struct DeviceOps {
  const DeviceRake *rakes;
  const DeviceCombine *combines;
  const DeviceAbsorb *absorbs;
  const DeviceCompress *compresses;
  const std::uint32_t *roots;
  const DeviceChain *chains;
  const DeviceMaxPiece *max_pieces;
};

struct Dims {
  std::uint32_t K;      // states
  std::uint32_t B;      // batch items in this call
  std::uint32_t E;      // edges (= shared slots)
  std::uint32_t N;      // nodes
  std::uint32_t roots;  // number of roots
};

// Observations of a categorical call, for the generic rake kernels to form
// the leaf vectors of the first `rakes` rake ops (whose leaves are original
// leaves) themselves instead of reading node state: the factor of node n,
// item b is (root potential if n is a root, else 1) times the emission row
// of the observed category (all zero for a category >= categories; 1 for an
// unobserved node), in the element type. rakes = 0: none.
struct CategoricalLeaves {
  const std::uint8_t *obs;        // [observation node][B]
  const std::uint32_t *index;     // [N] observation index of a node, 0xFFFFFFFF if unobserved
  const std::uint8_t *is_root;    // [N]
  const void *root, *emissions;   // [K], [categories][K]
  std::uint32_t categories;
  std::uint32_t rakes;
  std::uint32_t element_logs;     // logs in the element type (see LaunchInitCategorical)
};

// Scaled contraction state (element type float or double, erased).
//   nodes      [N][B][K] max-shifted logs     node_scale   [N][B]   FP64
//   branches   [R][B][K] max-shifted logs     branch_scale [R][B]   FP64
//   edges      [E][K][K] linear               edge_scale   [E][K]   FP64 row
//              scales, or null for unscaled caller potentials
//   paths      [C][B][K][K] linear, rows max-normalised
//                                             path_scale   [C][B][K] FP64
//   node_adj   [N][B][K]        branch_adj   [R][B][K]   (linear adjoints)
//   path_adj   [C][B][K][K]     edge_out     [B][E][K][K] (edge marginals)
struct LinearBuffers {
  void *nodes;
  double *node_scale;
  const void *edges;
  const double *edge_scale;
  void *paths;
  double *path_scale;
  void *branches;
  double *branch_scale;
  void *node_adj;
  void *branch_adj;
  void *path_adj;
  void *edge_out;
  // Log edge potentials [E][K][K] that the team rake kernels read in place
  // (exp of the row-shifted values on load) for edge slots; null when every
  // edge is staged in `edges`. The forward rake kernel writes each row's
  // shift (its maximum, 0 for a row of -inf) to log_edge_shift [E][K], which
  // the reverse reads.
  const void *log_edges;
  double *log_edge_shift;
  // Edge marginals (edge_out) stored as IEEE binary16 instead of the
  // workspace element type.
  int edge_out_half;
  // Edge slots hold FP32 values in an FP64 pass (the sampling state of an
  // FP32 workspace): the K = 8 and 16 kernels widen them on load, which is
  // exact. Only those kernels may run with it set.
  int edges_f32;
  // Leaf rakes of a categorical call (rake levels with cat.rakes != 0 run
  // on the generic rake kernels).
  CategoricalLeaves cat;
  // The caller's linear node potentials [B][N][K] (element type): set on the
  // launches of round 0's rakes (every leaf an original leaf) when the node
  // initialisation skipped those leaves; the team rake kernels then form the
  // leaf vectors as the initialisation would (logs in the element type,
  // shifted by the vector's maximum, FP64 scale).
  const void *raw_leaves;
  // Compressions read their edge slots from log_edges too (the K = 32 and 64
  // FP32 MFMA kernels; no other compression kernel may run with it set).
  // The forward one writes each row's shift (its maximum, -inf for a row of
  // -inf) to log_edge_shift, which must then alias edge_scale, and the
  // reverse reads it back through edge_scale. Otherwise compressions read
  // `edges`.
  int compress_log_edges;
};

// Log-domain max-product state (FP64):
//   nodes [N][B][K], edges [E][K][K], paths [C][B][K][K], branches [R][B][K],
//   rake_choice [R][B][K] u8, compress_choice [C][B][K][K] u8.
struct MaxBuffers {
  double *nodes;
  const double *edges;
  double *paths;
  double *branches;
  std::uint8_t *rake_choice;
  std::uint8_t *compress_choice;
  // Edge potentials [E][K][K] in the workspace precision (16-byte aligned)
  // that the K = 8 and 16 kernels log on load (x > 0 ? log x : -inf, as the
  // log transform into `edges` does); null when edges come from `edges`.
  const void *raw_edges;
  int raw_edges_f32;
  // Node potentials [B][N][K] in the workspace precision (16-byte aligned)
  // for round 0's first rake level and absorption levels, whose leaves and
  // parents MaxInit left unwritten: the team rake kernels log the leaf
  // values on load, and the absorption logs the parent's (the same
  // x > 0 ? log x : -inf); null when leaves and parents come from `nodes`.
  const void *raw_leaves;
  int raw_leaves_f32;
  // Per rake of the plan: 1 when its branch is absorbed as it stands (its
  // parent rakes no other child in that round). The team rake kernels then
  // add the rake's values into the parent's node vector (the absorption's
  // arithmetic) instead of writing the branch, and the absorption levels
  // hold the other absorptions only. Null: every rake writes its branch.
  const std::uint8_t *rake_direct;
};

enum class Real { kF32 = 0, kF64 = 1 };

// Launch variants chosen per workspace (defaults, thmm_set_tuning or the
// autotuner).
struct Tuning {
  std::uint32_t compress_items_forward = 2;  // K = 16 MFMA compressions per wavefront: 1 or 2
  std::uint32_t compress_items_reverse = 2;
  std::uint32_t rake_block = 256;            // threads per block of the team rake kernels: 64, 128, 256
  std::uint32_t gather_chunk = 16;           // rakes per team in the fused forward: 1 .. 64
  std::uint32_t compress_depth = 2;          // compression-tree levels per grouped compression: 1 (none) .. 3
  std::uint32_t rake_log_edges = 1;          // log-potential calls: rakes read log edges in place (1) or staged (0)
  std::uint32_t persistent = 0;              // K = 16 FP32 forests: each pass as one persistent launch (1)
  std::uint32_t input_leaves = 1;           // leaf rakes form their leaf vectors from the call's inputs (1)
};

struct Launch {
  Real real;
  Dims dims;
  DeviceOps ops;
  void *stream;  // hipStream_t
  Tuning tuning;
};

// ---- input staging ----
// Linear potentials, public [B][N][K] -> internal [N][B][K] max-shifted
// logs + FP64 scales. `in_real` is the element type of `input`; `nodes`
// has element type launch.real.
// With `only` (null: every node) just the listed nodes are written, for
// every item; a list needs the team transpose (FastTransposeFits and 16-byte
// aligned buffers).
struct NodeList {
  const std::uint32_t *ids;
  std::uint32_t count;
};
void LaunchInitNodes(const Launch &l, Real in_real, const void *input, void *nodes,
                     double *node_scale, const NodeList *only = nullptr);
// Log potentials [B][N][K] -> max-shifted logs + FP64 scales.
void LaunchInitNodesFromLogs(const Launch &l, const void *log_input, void *nodes,
                             double *node_scale);
// Log edge potentials [E][K][K] -> row-max-shifted exponentials + FP64 row
// scales [E][K].
// The same transform for the edges listed in ids only (edges and scales
// written in place); false when the state count has no team kernels (8, 16,
// 32, 64), in which case nothing is launched.
bool LaunchStageLogEdges(const Launch &l, const void *log_edges, const std::uint32_t *ids, std::uint32_t count,
                         void *edges, double *edge_scale);
void LaunchInitEdgesFromLogs(const Launch &l, const void *log_input, void *edges,
                             double *edge_scale);
// Observations [rows = batch][cols = observation nodes] (u8) -> [cols][rows],
// the layout LaunchInitCategorical reads.
void LaunchTransposeObservations(const Launch &l, const std::uint8_t *in, std::uint8_t *out, std::uint32_t rows,
                                 std::uint32_t cols);
// Categorical observations ([observation node][batch] u8) -> internal
// [N][B][K] node factors, the product
// root-or-one times emission formed in `product_real` so that it is bitwise
// equal to the dense construction. `raw` stores the products as they are
// (element type product_real); otherwise max-shifted logs of element type
// launch.real plus FP64 scales.
// Nodes with skip[n] set (null: none) are left unwritten. The logs are
// taken as the dense initialisation of the same factors takes them: in the
// element type when `element_logs` (an FP32 workspace whose dense
// initialisation runs on the team transpose, FastTransposeFits), else in
// FP64.
void LaunchInitCategorical(const Launch &l, Real product_real, bool raw,
                           const std::uint8_t *observations,
                           const std::uint32_t *observation_index_by_node,
                           std::uint32_t observation_count, std::uint32_t categories,
                           const std::uint8_t *is_root, const void *root_potential,
                           const void *emissions, void *nodes, double *node_scale,
                           const std::uint8_t *skip = nullptr, bool element_logs = false);
// Whether the initialisation of linear potentials [X][Y][K] (16-byte
// aligned) runs on the team transpose, which takes logs in the wider of the
// input and state element types rather than in FP64.
bool FastTransposeFits(std::uint32_t K, std::uint32_t X, std::uint32_t Y);
// sums[r] = sum over b of log_z[b * roots + r] (FP64), in a fixed order:
// 256 partial sums (partial t adds b = t, t + 256, ... in turn), then added
// pairwise (t and t + 128, then t and t + 64, ..., t and t + 1).
void LaunchSumOverBatch(const Launch &l, const double *log_z, double *sums);

// ---- fused rake + combination + absorption (sum-product forward) ----
// The rakes of one round, grouped by the absorption their branch reaches
// (the parent), in chunks of at most a few rakes: a chunk's messages are
// summed in FP64 and either absorbed into the parent at once (partial ==
// kDirectChunk) or written to partials[partial][b][K] for a segment pass.
constexpr std::uint32_t kDirectChunk = 0xFFFFFFFFu;
// leaf0/slot0 repeat the chunk's first rake, so that a one-rake chunk needs
// no second dependent load before its data.
struct DeviceGatherChunk {
  std::uint32_t first, count, parent, partial;
  std::uint32_t leaf0, slot0, pad0, pad1;
};
// Sums `count` consecutive partials into partial `out`, or, for out ==
// kDirectChunk, absorbs the sum into `parent`.
struct DeviceGatherSegment {
  std::uint32_t parent, first_partial, count, out;
};
struct GatherOps {
  const DeviceRake *rakes;  // sorted by target absorption within a round
  const DeviceGatherChunk *chunks;
  const DeviceGatherSegment *segments;
  double *partials;
};
// True when the fused kernels cover this state count and element type.
bool FusedForwardSupported(std::uint32_t states);
void LaunchGatherChunks(const Launch &l, const GatherOps &g, std::uint32_t offset, std::uint32_t count,
                        const LinearBuffers &buf);
void LaunchGatherSegments(const Launch &l, const GatherOps &g, std::uint32_t offset, std::uint32_t count,
                          const LinearBuffers &buf);

// ---- persistent passes (K = 16 FP32, batch 1) ----
// A pass's levels as one launch: resident blocks pull tasks (a run of ops of
// one level of one kind) from a queue in level order. The trees of the
// forest are cut into groups of consecutive trees; a task waits until every
// earlier level of its own group is done (the ops of one tree depend only on
// earlier levels of that tree), so groups run ahead of one another instead
// of meeting at every level boundary.
constexpr std::uint32_t kTaskGather = 0, kTaskSegments = 1, kTaskChainForward = 2, kTaskRakeReverse = 3,
                        kTaskChainReversePair = 4, kTaskChainReverseSmall = 5;
struct DeviceTask {
  std::uint32_t kind;
  std::uint32_t step;   // the group's level index among its levels in this pass
  std::uint32_t group;
  std::uint32_t offset, count;  // ops of the level's kind
  std::uint32_t blocks;         // block-sized runs of the ops, processed in turn
  std::uint32_t pad0, pad1;
};
struct PersistentQueue {
  const DeviceTask *tasks;
  std::uint32_t num_tasks;
  std::uint32_t steps, groups;  // (group, step) pairs; groups
  const std::uint32_t *need;         // [steps] tasks of each (group, step), group-major
  const std::uint32_t *need_offset;  // [groups] first (group, step) of each group
  // [1 + steps + groups]: the queue head, finished tasks per (group, step)
  // and completed steps per group; cleared by the launch.
  std::uint32_t *counters;
  // Set (and never cleared by a launch) when a wait exceeds its spin limit.
  std::uint32_t *failure;
  std::uint32_t grid;  // blocks launched: at most the blocks the device holds at once
  std::uint64_t wait_ticks;  // longest dependency wait, in ticks of the device's constant wall clock
};
// Ticks of the device's constant wall clock in `seconds`.
std::uint64_t WallClockTicks(double seconds);
// Blocks of the pass's kernels the device holds at once (0 if unknown).
std::uint32_t PersistentGrid(bool reverse);
void LaunchPersistentPass(const Launch &l, bool reverse, const GatherOps &g, const PersistentQueue &q,
                          const LinearBuffers &buf);

// ---- scaled forward (log Z) ----
void LaunchForwardLevel(const Launch &l, std::uint32_t primitive, std::uint32_t offset,
                        std::uint32_t count, const LinearBuffers &buf);
// log_z[b * roots + r] (FP64).
void LaunchFinish(const Launch &l, const LinearBuffers &buf, double *log_z);

// ---- reverse (adjoint = marginals) ----
// node_adj(root) = root / sum(root), scaled by grad[b * roots + r] if given.
void LaunchSeedAdjoint(const Launch &l, const LinearBuffers &buf, const double *grad);
void LaunchReverseLevel(const Launch &l, std::uint32_t primitive, std::uint32_t offset,
                        std::uint32_t count, const LinearBuffers &buf);
// node_adj [N][B][K] -> out [B][N][K].
void LaunchNodeAdjointOut(const Launch &l, const void *node_adj, void *out);
// out[e][.] = sum_b edge_out[b][e][.]
void LaunchReduceEdgesOverBatch(const Launch &l, const void *edge_out, void *out);
// Branch-length chain for edge factors P_e = expm(Q t_e), dP_e/dt_e = Q P_e:
//   grad[e] = sum_{p,c} M_e[p][c] (Q P_e)[p][c] / P_e[p][c],
// M_e = sum_b edge_out[b][e] (the adjoint of P_e, batch-summed), FP64.
// `terms` is [E][K][K] FP64 scratch; entries with P_e[p][c] = 0 contribute 0.
// edges[e][i][j] = max(0, sum_k V[i][k] exp(lambda_k t_e) V_inv[k][j]) in
// FP64, rounded to launch.real: expm(Q t_e) from the eigendecomposition of a
// reversible rate matrix (factors [K][K], [K], [K][K] and t [E], FP64).
void LaunchExpmEdges(const Launch &l, const double *V, const double *lambda, const double *V_inv,
                     const double *branch_lengths, void *edges);
void LaunchBranchLengthChain(const Launch &l, const void *edge_out, const void *edges, const double *rate,
                             double *terms, double *grad);
// Widens `count` FP32 values to FP64.
void LaunchWidenF32(const Launch &l, const float *in, double *out, std::size_t count);
// Casts `count` FP64 values to the launch element type.
void LaunchCastF64(const Launch &l, const double *in, void *out, std::size_t count);

// ---- max-product (FP64) ----
// With `only` (null: every node) just the listed nodes are written, for
// every item.
void LaunchMaxInit(const Launch &l, Real in_real, const void *nodes_in, const void *edges_in,
                   double *nodes, double *edges,
                   const NodeList *only = nullptr);
void LaunchMaxLevel(const Launch &l, std::uint32_t primitive, std::uint32_t offset,
                    std::uint32_t count, const MaxBuffers &buf);
void LaunchMaxFinish(const Launch &l, const MaxBuffers &buf, double *log_weight,
                     std::uint32_t *assignment);
void LaunchMaxExpandLevel(const Launch &l, std::uint32_t primitive, std::uint32_t offset,
                          std::uint32_t count, const MaxBuffers &buf,
                          std::uint32_t *assignment);

// ---- posterior sampling (FP64 scaled state) ----
// uniforms [B][N] (element type in_real), assignment [B][N]. A root with no
// supported state sets *failure.
// Sets one device word to zero with a kernel (not a memset node when a
// caller captures the stream into a graph).
void LaunchClearWord(const Launch &l, std::uint32_t *word);
void LaunchSampleSeed(const Launch &l, const LinearBuffers &buf, Real in_real,
                      const void *uniforms, std::uint32_t *assignment, std::uint32_t *failure);
void LaunchSampleExpandLevel(const Launch &l, std::uint32_t primitive, std::uint32_t offset,
                             std::uint32_t count, const LinearBuffers &buf, Real in_real,
                             const void *uniforms, std::uint32_t *assignment,
                             std::uint32_t *failure);

// Group width (lanes per batch item) used for a state count.
std::uint32_t GroupWidth(std::uint32_t states);

// Reports the first launch error of the calling thread and clears it; empty
// when none. The executor turns a non-empty result into runtime_error.
const char *LastLaunchError();

}  // namespace thmm

#endif  // TREE_HMM_AMD_LAUNCH_H_
