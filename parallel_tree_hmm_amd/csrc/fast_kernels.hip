// Coalesced and MFMA kernels for the hot levels (see kernels.hip for the
// number representation and the generic forms these replace).
//
// Team layout (rakes, K in {8, 16, 32, 64}): a team of T lanes owns one item.
// The item's K x K matrix is read and written as row-major 16-byte quads:
// in step j lane l moves quad j*T + l, i.e. row j*RPS + l/LPR, columns
// (l%LPR)*4 .. +3, so every memory instruction of a team touches one
// contiguous T*16-byte span (LPR = K/4 lanes per row, RPS = 4T/K rows per
// step). Row sums reduce over the LPR lanes of a row, column sums over the
// lanes that share columns.
//
// Fused rakes (K in {8, 16, 32, 64}): the same team walks a chunk of rakes that
// share a parent and folds their summed messages into the parent, replacing
// the combination and absorption levels of the sum-product forward pass.
//
// MFMA compression (FP32, K = 16, 32, 64): operands loaded straight from HBM
// in the layouts v_mfma_f32_16x16x4_f32 consumes (see the sections below;
// two items per wavefront at K = 16, one at K = 32 and 64); the reverse pass
// transposes one K x K tile through LDS.
// Every MFMA operand lies in [0, 1]: L and R are row-normalised, the middle
// weights are exp(w - max w), and in the reverse pass the ratio matrix adj/S
// is row-normalised with its row scales (and their maximum) kept in FP64.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.h"
#include "launch.h"

namespace thmm {
namespace {

using namespace dev;

// Cross-lane maxima and exchanges without an LDS round trip: DPP moves
// inside a 16-lane row (quad permutations, then row rotations by 4 and 8
// cover all 16 lanes), and ds_swizzle's bitmask mode for lane ^ 16 inside a
// 32-lane half. No source lane is ever out of its row, so every lane reads
// a valid value.
constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E, kRowHalfMirror = 0x141, kRowRotate4 = 0x124, kRowRotate8 = 0x128;
constexpr int kSwizzleXor16 = 0x401F;  // and_mask 0x1F, or_mask 0, xor_mask 0x10

template <int kCtrl> __device__ inline float DppMove(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xF, 0xF, false));
}
template <int kCtrl> __device__ inline double DppMove(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(bits), kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(bits >> 32), kCtrl, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Maximum over the 16 lanes of each row (lane l and l ^ 1, ..., l ^ 8).
__device__ inline float RowLanesMax(float v) {
  v = fmaxf(v, DppMove<kQuadXor1>(v));
  v = fmaxf(v, DppMove<kQuadXor2>(v));
  v = fmaxf(v, DppMove<kRowRotate4>(v));
  return fmaxf(v, DppMove<kRowRotate8>(v));
}
__device__ inline double RowLanesMax(double v) {
  v = fmax(v, DppMove<kQuadXor1>(v));
  v = fmax(v, DppMove<kQuadXor2>(v));
  v = fmax(v, DppMove<kRowRotate4>(v));
  return fmax(v, DppMove<kRowRotate8>(v));
}

// The value of lane l ^ 16.
__device__ inline float Xor16(float v) {
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), kSwizzleXor16));
}


// Maximum over each group of kLpr consecutive lanes (kLpr = 2, 4, 8, 16,
// aligned groups inside a 16-lane row): quad permutations, then the mirror
// of each 8-lane half (lane i <-> 7 - i) or rotations by 4 and 8.
template <int kLpr, typename V> __device__ inline V GroupLanesMax(V v) {
  static_assert(kLpr == 1 || kLpr == 2 || kLpr == 4 || kLpr == 8 || kLpr == 16, "lane group");
  if constexpr (kLpr >= 2) v = Max(v, DppMove<kQuadXor1>(v));
  if constexpr (kLpr >= 4) v = Max(v, DppMove<kQuadXor2>(v));
  if constexpr (kLpr == 8) v = Max(v, DppMove<kRowHalfMirror>(v));
  if constexpr (kLpr == 16) {
    v = Max(v, DppMove<kRowRotate4>(v));
    v = Max(v, DppMove<kRowRotate8>(v));
  }
  return v;
}

// Log edge potentials read in place by the rake kernels: each row of the
// team's quads becomes exp(x - shift), the shift being the row's maximum M,
// or 0 for a row of -inf (which then stays 0); the values equal those of
// the log-edge staging transform (ExpNonPositive) bit for bit. Returns M of
// step `owned`'s row (-inf for a dead row).
template <typename T> __device__ inline T RowShift(T top) { return top > NegInf<T>() ? top : T(0); }

template <typename T, int kLpr, int kJ>
__device__ inline double ExpLogRows(Quad<T> (&pq)[kJ], int owned) {
  double own = 0.0;
#pragma unroll
  for (int j = 0; j < kJ; ++j) {
    T top = Max(Max(pq[j].v[0], pq[j].v[1]), Max(pq[j].v[2], pq[j].v[3]));
    top = GroupLanesMax<kLpr>(top);
    const T shift = RowShift(top);
#pragma unroll
    for (int i = 0; i < 4; ++i) pq[j].v[i] = ExpNonPositive(static_cast<T>(pq[j].v[i] - shift));
    if (j == owned) own = top > NegInf<T>() ? static_cast<double>(top) : kNegInf;
  }
  return own;
}

template <int K> struct Team {
  static constexpr int kLanesPerRow = K / 4;
  static constexpr int kLanes = K <= 4 ? 4 : (K <= 16 ? 16 : 64);
  static constexpr int kRowsPerStep = 4 * kLanes / K;
  static constexpr int kSteps = K / kRowsPerStep;
};
static_assert(Team<8>::kSteps == 1 && Team<16>::kSteps == 4 && Team<32>::kSteps == 4 && Team<64>::kSteps == 16,
              "team tiling");
// One logarithm per lane in RakeForwardTeam needs at most one step per lane
// of a row.
static_assert(Team<8>::kSteps <= Team<8>::kLanesPerRow && Team<16>::kSteps <= Team<16>::kLanesPerRow &&
                  Team<32>::kSteps <= Team<32>::kLanesPerRow && Team<64>::kSteps <= Team<64>::kLanesPerRow,
              "row ownership");

// A team's rows are owned one per lane: row j kRps + l / kLpr on the lane
// with l % kLpr == j % kLpr. RowToLane moves row r's value to lane r (r <
// K; lanes from K on get some row's value) so that a store of the team's
// rows runs in lane order: a store whose lanes are scattered over a line
// costs one partial-line write to HBM per lane.
template <int K, typename V> __device__ inline V RowToLane(V value, int l) {
  using TM = Team<K>;
  constexpr int kLpr = TM::kLanesPerRow, kRps = TM::kRowsPerStep;
  const int r = l % K;
  return __shfl(value, (r % kRps) * kLpr + r / kRps, TM::kLanes);
}

// log v for v > 0, else -inf (NaN included), in the element type: the
// node initialisation's log of a linear potential.
template <typename T> __device__ inline T LogOf(T v) { return v > T(0) ? log(v) : NegInf<T>(); }
template <> __device__ inline float LogOf(float v) { return v > 0.0f ? logf(v) : NegInf<float>(); }

// exp of the team's quad of a K-vector: the lanes sharing a quad (same
// l % kLpr) each exponentiate one element and exchange the results, so every
// element is exponentiated once per team.
template <typename T, int kT, int kLpr>
__device__ inline void TeamExp(const Quad<T> &v, int l, T (&e)[4]) {
  static_assert(kT / kLpr >= 4, "a quad needs four lanes");
  const T mine = Exp(v.v[(l / kLpr) % 4]);
#pragma unroll
  for (int i = 0; i < 4; ++i) e[i] = __shfl(mine, l % kLpr + kLpr * i, kT);
}

// A rake's leaf vector: the node state's max-shifted logs, or (kRaw) the
// caller's linear potentials [B][N][K] (buf.raw_leaves).
template <typename T, int K, bool kRaw>
__device__ inline const T *LeafVector(const LinearBuffers &buf, const Dims &d, std::uint32_t leaf, std::uint32_t b) {
  if constexpr (kRaw)
    return static_cast<const T *>(buf.raw_leaves) + (static_cast<std::uint64_t>(b) * d.N + leaf) * K;
  return static_cast<const T *>(buf.nodes) + (static_cast<std::uint64_t>(leaf) * d.B + b) * K;
}

// TeamExp of a leaf vector given as linear potentials (x: the lane's quad),
// formed as the node initialisation (TransposeTeam, kMode 1, same element
// type) forms the vector: w = log x in the element type, top = the vector's
// maximum, v = w - top (-inf without support); returns the FP64 scale (top,
// or -inf). Each lane logs the one element of its quad that it
// exponentiates, and the team's maximum covers the whole vector.
template <typename T, int kT, int kLpr>
__device__ inline double RawLeafExp(const Quad<T> &x, int l, T (&e)[4]) {
  const T w = LogOf(x.v[(l / kLpr) % 4]);
  T top = w;
#pragma unroll
  for (int o = kT / 2; o > 0; o >>= 1) top = Max(top, __shfl_xor(top, o, kT));
  const bool live = top > NegInf<T>();
  const T mine = Exp(live ? static_cast<T>(w - top) : NegInf<T>());
#pragma unroll
  for (int i = 0; i < 4; ++i) e[i] = __shfl(mine, l % kLpr + kLpr * i, kT);
  return live ? static_cast<double>(top) : kNegInf;
}

// log of a positive value: FP32 mantissa logarithm plus the exact exponent
// term in FP64 (absolute error about 6e-8 whatever the magnitude).
__device__ inline double LogPositive(float v) {
  int exponent = 0;
  const float mantissa = frexpf(v, &exponent);
  return static_cast<double>(logf(mantissa)) + 0.6931471805599453 * exponent;
}
__device__ inline double LogPositive(double v) { return log(v); }

// branch[p] = log sum_c path[p][c] exp(leaf[c]) (src/inference.cc:704-723)
template <typename T, int K>
__global__ __launch_bounds__(256) void RakeForwardTeam(DeviceOps ops, Dims d, std::uint32_t offset,
                                                       std::uint32_t count, LinearBuffers buf) {
  using TM = Team<K>;
  constexpr int kT = TM::kLanes, kLpr = TM::kLanesPerRow, kRps = TM::kRowsPerStep, kJ = TM::kSteps;
  const std::uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const std::uint32_t item = tid / kT;
  const int l = static_cast<int>(tid % kT);
  if (item >= count * d.B) return;
  const std::uint32_t b = item % d.B;
  const DeviceRake op = ops.rakes[offset + item / d.B];
  const std::uint64_t leaf_at = static_cast<std::uint64_t>(op.leaf) * d.B + b;
  const int c0 = (l % kLpr) * 4;
  const Quad<T> vq = LoadQuad(static_cast<const T *>(buf.nodes) + leaf_at * K + c0);
  T e[4];
  TeamExp<T, kT, kLpr>(vq, l, e);
  // After the row reduction all kLpr lanes of a row hold its sum; the lane
  // with l % kLpr == j % kLpr takes the logarithm, so each row's log is
  // formed once; RowToLane puts the rows in lane order for the store.
  T sums[kJ];
#pragma unroll
  for (int j = 0; j < kJ; ++j) {
    const Quad<T> p = LoadSlotQuad<T>(buf, d, op.slot, b, static_cast<std::uint32_t>(j * kT + l) * 4);
    T part = p.v[0] * e[0] + p.v[1] * e[1] + p.v[2] * e[2] + p.v[3] * e[3];
#pragma unroll
    for (int o = kLpr / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, kT);
    sums[j] = part;
  }
  int row = -1;
  T part = T(0);
#pragma unroll
  for (int j = 0; j < kJ; ++j)
    if (j % kLpr == l % kLpr) row = j * kRps + l / kLpr, part = sums[j];
  const double message = row >= 0 && part > T(0) ? RowScale(buf, d, op.slot, b, row) + LogPositive(part) : kNegInf;
  double top = message;
#pragma unroll
  for (int o = kT / 2; o > 0; o >>= 1) top = fmax(top, __shfl_xor(top, o, kT));
  const std::uint64_t at = static_cast<std::uint64_t>(op.branch) * d.B + b;
  const bool live = top > kNegInf;
  const double ordered = RowToLane<K>(message, l);
  if (l < K) static_cast<T *>(buf.branches)[at * K + l] = live ? static_cast<T>(ordered - top) : NegInf<T>();
  if (l == 0) buf.branch_scale[at] = live ? buf.node_scale[leaf_at] + top : kNegInf;
}

// Rake adjoint (src/inference.cc:488-518): with t = P[p][c] exp(v[c]) and
// m[p] = sum_c t, adj_P = adj_m[p] t / m[p] (ratio in FP64) and
// adj_v[c] = sum_p adj_P[p][c]. kLogEdges: buf.log_edges may be set (the
// variant without it holds fewer registers).
// `tid` is the lane's index in the launch (block * block size + thread).
// kRaw: the leaves are original leaves read from buf.raw_leaves.
template <typename T, int K, bool kLogEdges, bool kRaw = false>
__device__ __forceinline__ void RakeReverseTeamBody(const DeviceOps &ops, const Dims &d, std::uint32_t offset,
                                                    std::uint32_t count, const LinearBuffers &buf,
                                                    std::uint32_t tid) {
  using TM = Team<K>;
  constexpr int kT = TM::kLanes, kLpr = TM::kLanesPerRow, kRps = TM::kRowsPerStep, kJ = TM::kSteps;
  const std::uint32_t item = tid / kT;
  const int l = static_cast<int>(tid % kT);
  if (item >= count * d.B) return;
  const std::uint32_t b = item % d.B;
  const DeviceRake op = ops.rakes[offset + item / d.B];
  const std::uint64_t leaf_at = static_cast<std::uint64_t>(op.leaf) * d.B + b;
  const int c0 = (l % kLpr) * 4;
  const Quad<T> vq = LoadQuad(LeafVector<T, K, kRaw>(buf, d, op.leaf, b) + c0);
  T e[4];
  if constexpr (kRaw)
    RawLeafExp<T, kT, kLpr>(vq, l, e);
  else
    TeamExp<T, kT, kLpr>(vq, l, e);
  // Log edges (see RakeGatherTeam): exp(x - shift) with the row shifts the
  // forward pass stored, the values the forward pass used.
  const bool log_edge = kLogEdges && buf.log_edges != nullptr && op.slot < d.E;
  const T *P = log_edge ? static_cast<const T *>(buf.log_edges) + static_cast<std::uint64_t>(op.slot) * K * K
                        : Slot<T>(buf, d, op.slot, b);
  const AdjointTarget<T> adj_p = SlotAdjointTarget<T>(buf, d, op.slot, b);
  const T *adj_m = RakeAdjoint<T>(buf, d, op, b);
  double column[4] = {0.0, 0.0, 0.0, 0.0};
  // FP32: a step's ratios a / part are taken with the hardware reciprocal
  // and its column sums kept in FP32 when, on every lane, part is 0 or at
  // least 2^-100 and |a| below 2^26 (so a / part stays below 2^126), else
  // (wave-uniform) in FP64.
  T narrow[4] = {T(0), T(0), T(0), T(0)};
  constexpr bool kFloat = std::is_same<T, float>::value;
  // Four steps unrolled at a time: at K = 64 a full unroll holds all 16
  // steps' loads in registers (three wavefronts per SIMD).
#pragma unroll 4
  for (int j = 0; j < kJ; ++j) {
    const std::uint32_t quad = static_cast<std::uint32_t>(j * kT + l) * 4;
    Quad<T> p = LoadQuadStream(P + quad);
    if (log_edge) {
      const T shift =
          static_cast<T>(buf.log_edge_shift[static_cast<std::uint64_t>(op.slot) * K + j * kRps + l / kLpr]);
#pragma unroll
      for (int i = 0; i < 4; ++i) p.v[i] = ExpNonPositive(static_cast<T>(p.v[i] - shift));
    }
    T t[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = p.v[i] * e[i];
    T part = t[0] + t[1] + t[2] + t[3];
#pragma unroll
    for (int o = kLpr / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, kT);
    const T a = adj_m[j * kRps + l / kLpr];
    Quad<T> out;
    if (kFloat && __all(!(part > T(0) && part < T(0x1p-100f)) && fabs(a) < T(0x1p26f))) {
      const T g = (a != T(0) && part > T(0)) ? a * static_cast<T>(__builtin_amdgcn_rcpf(static_cast<float>(part)))
                                             : T(0);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        out.v[i] = g * t[i];
        narrow[i] += out.v[i];
      }
    } else {
      const double g = (a != T(0) && part > T(0)) ? static_cast<double>(a) / static_cast<double>(part) : 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double c = g * static_cast<double>(t[i]);
        out.v[i] = static_cast<T>(c);
        column[i] += c;
      }
    }
    StoreAdjointQuad(adj_p, quad, out);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) column[i] += static_cast<double>(narrow[i]);
#pragma unroll
  for (int o = kLpr; o < kT; o <<= 1)
#pragma unroll
    for (int i = 0; i < 4; ++i) column[i] += __shfl_xor(column[i], o, kT);
  if (l < kLpr) {
    Quad<T> q;
#pragma unroll
    for (int i = 0; i < 4; ++i) q.v[i] = static_cast<T>(column[i]);
    StoreQuad(static_cast<T *>(buf.node_adj) + leaf_at * K + c0, q);
  }
}

template <typename T, int K, bool kLogEdges, bool kRaw = false>
__global__ __launch_bounds__(256) void RakeReverseTeam(DeviceOps ops, Dims d, std::uint32_t offset,
                                                       std::uint32_t count, LinearBuffers buf) {
  RakeReverseTeamBody<T, K, kLogEdges, kRaw>(ops, d, offset, count, buf, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Fused rake + combination + absorption
// ---------------------------------------------------------------------------
//
// Combining branches multiplies messages and absorbing multiplies the product
// into the parent (src/inference.cc:173-189, 725-747): in logs, the parent's
// new vector is its old one plus the sum of its raked children's messages. A
// team walks one chunk of rakes that share a parent, accumulates their
// absolute log messages in FP64 (lane-owned rows, as in RakeForwardTeam) and
// either folds the sum into the parent or leaves it as a partial.

// Adds the FP64 absolute log vector `acc` (rows owned by lanes with row >= 0)
// into node vector `at` (node `parent`, item `b`) and re-shifts it by its
// maximum. kRaw: the parent absorbs for the first time in round 0 and the
// node initialisation skipped it; its vector is formed here from the caller's
// linear potentials [B][N][K] (buf.raw_leaves) as TransposeTeam (kMode 1,
// same element type) forms it: w = log x, top = the maximum over the vector
// (exact, so the reduction order does not matter), w - top in the element
// type and top as the FP64 scale.
template <typename T, int K, int kT, bool kRaw = false>
__device__ inline void AbsorbSum(const LinearBuffers &buf, const Dims &d, std::uint32_t parent, std::uint32_t b,
                                 int row, double acc, int lane) {
  const std::uint64_t at = static_cast<std::uint64_t>(parent) * d.B + b;
  T *x = static_cast<T *>(buf.nodes) + at * K;
  T value;
  double scale;
  if constexpr (kRaw) {
    const T *raw = static_cast<const T *>(buf.raw_leaves) + (static_cast<std::uint64_t>(b) * d.N + parent) * K;
    const T w = row >= 0 ? LogOf(raw[row]) : NegInf<T>();
    T vector_top = w;
#pragma unroll
    for (int o = kT / 2; o > 0; o >>= 1) vector_top = Max(vector_top, __shfl_xor(vector_top, o, kT));
    const bool support = vector_top > NegInf<T>();
    value = support ? static_cast<T>(w - vector_top) : NegInf<T>();
    scale = support ? static_cast<double>(vector_top) : kNegInf;
  } else {
    value = row >= 0 ? x[row] : T(0);
    scale = buf.node_scale[at];
  }
  const double y = row >= 0 ? static_cast<double>(value) + scale + acc : kNegInf;
  double top = y;
#pragma unroll
  for (int o = kT / 2; o > 0; o >>= 1) top = fmax(top, __shfl_xor(top, o, kT));
  const bool live = top > kNegInf;
  if (row >= 0) x[row] = live ? static_cast<T>(y - top) : NegInf<T>();
  if (lane == 0) buf.node_scale[at] = live ? top : kNegInf;
}

// kLogEdges: buf.log_edges may be set (the variant without it holds fewer
// registers). FP32 edge slots of an FP64 pass (buf.edges_f32) are widened
// as they load, which is exact. kRaw: round 0's chunks; the leaves are
// original leaves and the direct chunks' parents absorb for the first time,
// both read from buf.raw_leaves.
template <typename T, int K, bool kLogEdges, bool kRaw = false>
__device__ __forceinline__ void RakeGatherTeamBody(const GatherOps &g, const Dims &d, std::uint32_t offset,
                                                   std::uint32_t count, const LinearBuffers &buf, std::uint32_t tid) {
  using TM = Team<K>;
  constexpr int kT = TM::kLanes, kLpr = TM::kLanesPerRow, kRps = TM::kRowsPerStep, kJ = TM::kSteps;
  const std::uint32_t item = tid / kT;
  const int l = static_cast<int>(tid % kT);
  if (item >= count * d.B) return;
  const std::uint32_t b = item % d.B;
  const DeviceGatherChunk chunk = g.chunks[offset + item / d.B];
  int owned = -1;
#pragma unroll
  for (int j = 0; j < kJ; ++j)
    if (j % kLpr == l % kLpr) owned = j;
  const int row = owned >= 0 ? owned * kRps + l / kLpr : -1;
  const int c0 = (l % kLpr) * 4;
  double acc = 0.0;
  DeviceRake op{chunk.leaf0, chunk.slot0, 0, 0};
  for (std::uint32_t i = 0; i < chunk.count; ++i) {
    DeviceRake next = op;
    if (i + 1 < chunk.count) next = g.rakes[chunk.first + i + 1];
    const std::uint64_t leaf_at = static_cast<std::uint64_t>(op.leaf) * d.B + b;
    const Quad<T> vq = LoadQuad(LeafVector<T, K, kRaw>(buf, d, op.leaf, b) + c0);
    const bool log_edge = kLogEdges && buf.log_edges != nullptr && op.slot < d.E;  // the same for the whole team
    const T *P = log_edge ? static_cast<const T *>(buf.log_edges) + static_cast<std::uint64_t>(op.slot) * K * K
                          : Slot<T>(buf, d, op.slot, b);
    Quad<T> pq[kJ];
    bool widen = false;
    if constexpr (std::is_same<T, double>::value) widen = buf.edges_f32 != 0 && op.slot < d.E;
    if (widen) {
      const float *F = static_cast<const float *>(buf.edges) + static_cast<std::uint64_t>(op.slot) * K * K;
#pragma unroll
      for (int j = 0; j < kJ; ++j) {
        const Quad<float> f = LoadQuadStream(F + static_cast<std::uint32_t>(j * kT + l) * 4);
#pragma unroll
        for (int i = 0; i < 4; ++i) pq[j].v[i] = static_cast<T>(f.v[i]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < kJ; ++j) pq[j] = LoadQuadStream(P + static_cast<std::uint32_t>(j * kT + l) * 4);
    }
    double leaf_scale = 0.0;
    if constexpr (!kRaw) leaf_scale = buf.node_scale[leaf_at];
    double row_scale;
    if (log_edge) {
      row_scale = ExpLogRows<T, kLpr, kJ>(pq, owned);
      const double ordered = RowToLane<K>(row_scale, l);
      if (l < K)
        buf.log_edge_shift[static_cast<std::uint64_t>(op.slot) * K + l] = ordered > kNegInf ? ordered : 0.0;
    } else {
      row_scale = row >= 0 ? RowScale(buf, d, op.slot, b, row) : 0.0;
    }
    T e[4];
    if constexpr (kRaw)
      leaf_scale = RawLeafExp<T, kT, kLpr>(vq, l, e);
    else
      TeamExp<T, kT, kLpr>(vq, l, e);
    T mine = T(0);
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
      T part = pq[j].v[0] * e[0] + pq[j].v[1] * e[1] + pq[j].v[2] * e[2] + pq[j].v[3] * e[3];
#pragma unroll
      for (int o = kLpr / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, kT);
      if (j == owned) mine = part;
    }
    acc += (row >= 0 && mine > T(0)) ? leaf_scale + row_scale + LogPositive(mine) : kNegInf;
    op = next;
  }
  // Rows in lane order for the stores below.
  const double ordered = RowToLane<K>(acc, l);
  const int lane_row = l < K ? l : -1;
  if (chunk.partial == kDirectChunk)
    AbsorbSum<T, K, kT, kRaw>(buf, d, chunk.parent, b, lane_row, ordered, l);
  else if (l < K)
    g.partials[(static_cast<std::uint64_t>(chunk.partial) * d.B + b) * K + l] = ordered;
}

template <typename T, int K, bool kLogEdges, bool kRaw = false>
__global__ __launch_bounds__(256) void RakeGatherTeam(GatherOps g, Dims d, std::uint32_t offset, std::uint32_t count,
                                                      LinearBuffers buf) {
  RakeGatherTeamBody<T, K, kLogEdges, kRaw>(g, d, offset, count, buf, blockIdx.x * blockDim.x + threadIdx.x);
}

// Sums a run of partials (K lanes, lane = row) into a partial of the next
// reduction level or, at the last level, into the parent. kRaw: round 0's
// reduction levels, whose parents the node initialisation skipped.
template <typename T, int K, bool kRaw = false>
__device__ __forceinline__ void GatherSegmentsTeamBody(const GatherOps &g, const Dims &d, std::uint32_t offset,
                                                       std::uint32_t count, const LinearBuffers &buf,
                                                       std::uint32_t tid) {
  const std::uint32_t item = tid / K;
  const int row = static_cast<int>(tid % K);
  if (item >= count * d.B) return;
  const std::uint32_t b = item % d.B;
  const DeviceGatherSegment seg = g.segments[offset + item / d.B];
  double acc = 0.0;
  for (std::uint32_t i = 0; i < seg.count; ++i)
    acc += g.partials[(static_cast<std::uint64_t>(seg.first_partial + i) * d.B + b) * K + row];
  if (seg.out == kDirectChunk)
    AbsorbSum<T, K, K, kRaw>(buf, d, seg.parent, b, row, acc, row);
  else
    g.partials[(static_cast<std::uint64_t>(seg.out) * d.B + b) * K + row] = acc;
}

template <typename T, int K, bool kRaw = false>
__global__ __launch_bounds__(256) void GatherSegmentsTeam(GatherOps g, Dims d, std::uint32_t offset,
                                                          std::uint32_t count, LinearBuffers buf) {
  GatherSegmentsTeamBody<T, K, kRaw>(g, d, offset, count, buf, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// FP32 MFMA compression, K = 16
// ---------------------------------------------------------------------------
//
// v_mfma_f32_16x16x4_f32 accumulates D += A B over four k per instruction;
// lane l supplies A[l%16][k] and B[k][l%16] for k = 4(l/16) + kb in step kb
// (the k order within a product is free), and D[i][j] lands in register i%4
// of lane j + 16(i/4). Two per-lane layouts of a 16 x 16 matrix X follow:
//   column layout: lane l, register r holds X[4(l/16) + r][l%16]
//                  (four 4-byte loads, each a 64-byte segment per 16 lanes);
//   row layout:    lane l, register r holds X[l%16][4(l/16) + r]
//                  (one 16-byte load per lane).
// D comes out in column layout; a matrix in column layout is a B operand as
// it stands and an A operand transposed, and the row layout of X is the
// column layout of X^T. Hence S^T = R^T (L diag(e))^T takes R in column
// layout and L in row layout and yields S in row layout, so rows of S are
// reduced across 4 lanes and stored with 16-byte stores.

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;  // wavefronts per block

__device__ inline floatx4 Mfma(const float (&a)[4], const float (&b)[4]) {
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[kb], b[kb], acc, 0, 0, 0);
  return acc;
}

// The multiplier that normalises a row of values x <= top by its maximum
// top (0 for a dead row). A reciprocal of a top below 2^-100 can overflow
// or be flushed, so when any live row of the wavefront has one, the
// wavefront first scales its values (scale_values(2^100)) and tops by 2^100,
// which is exact and leaves every normalised value unchanged. Products here
// are at most K, so the scaled values stay finite.
template <typename F>
__device__ inline float RowInverse(float top, bool live, F &&scale_values) {
  if (__any(live && top < 0x1p-100f)) {
    scale_values(0x1p100f);
    top *= 0x1p100f;
  }
  return live ? __builtin_amdgcn_rcpf(top) : 0.0f;
}

// kStream: the kernel's last read of these values (non-temporal loads).
template <bool kStream = false> __device__ inline void LoadColumns(const float *X, int lane, float (&out)[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const float *at = X + (4 * (lane / 16) + r) * 16 + lane % 16;
    out[r] = kStream ? __builtin_nontemporal_load(at) : *at;
  }
}

template <bool kStream = false> __device__ inline void LoadRows(const float *X, int lane, float (&out)[4]) {
  const float *at = X + (lane % 16) * 16 + 4 * (lane / 16);
  const Quad<float> q = kStream ? LoadQuadStream(at) : LoadQuad(at);
#pragma unroll
  for (int r = 0; r < 4; ++r) out[r] = q.v[r];
}

// Operands of one compression as seen by one lane. An item is handled by a
// whole wavefront, so the op and slot bases are made wave-uniform
// (readfirstlane): their arithmetic runs on the scalar unit and the vector
// loads add only the lane's offset.
struct CompressOperands {
  DeviceCompress op;
  std::uint32_t b;
  const float *left, *right;
  float left_rows[4];       // L, row layout
  float right_cols[4];      // R, column layout
  float mid_own;            // mid[l%16]
  double right_scale_own;   // row scale of R at m = l%16
};

__device__ inline std::uint32_t Uniform(std::uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

template <bool kStream>
__device__ inline void LoadCompression(const DeviceOps &ops, const Dims &d, const LinearBuffers &buf,
                                       std::uint32_t offset, std::uint32_t item, int lane, CompressOperands &o) {
  item = Uniform(item);
  o.b = item % d.B;
  const DeviceCompress op = ops.compresses[offset + item / d.B];
  o.op.left = Uniform(op.left), o.op.middle = Uniform(op.middle), o.op.right = Uniform(op.right);
  o.op.out = Uniform(op.out);
  o.left = Slot<float>(buf, d, o.op.left, o.b);
  o.right = Slot<float>(buf, d, o.op.right, o.b);
  LoadRows<kStream>(o.left, lane, o.left_rows);
  LoadColumns<kStream>(o.right, lane, o.right_cols);
  const std::uint64_t mid_at = static_cast<std::uint64_t>(o.op.middle) * d.B + o.b;
  o.mid_own = static_cast<const float *>(buf.nodes)[mid_at * 16 + lane % 16];
  o.right_scale_own = RowScale(buf, d, o.op.right, o.b, lane % 16);
}

// Middle weights e[m] = exp(mid[m] + r_m - W): each lane forms the weight of
// m = l%16 (e_own) and gathers those of m = 4(l/16) + r into e[r]; returns
// W = max_m (mid[m] + r_m) (-inf when no middle state is live).
__device__ inline double StepWeights16(float mid_own, double right_scale_own, int lane, float (&e)[4],
                                      float *e_own) {
  const double w = static_cast<double>(mid_own) + right_scale_own;
  const double top = RowLanesMax(w);
  *e_own = top > kNegInf ? ExpNonPositive(static_cast<float>(w - top)) : 0.0f;
#pragma unroll
  for (int r = 0; r < 4; ++r) e[r] = __shfl(*e_own, 4 * (lane / 16) + r, 64);
  return top;
}

__device__ inline double MiddleWeights16(const CompressOperands &o, int lane, float (&e)[4], float *e_own) {
  return StepWeights16(o.mid_own, o.right_scale_own, lane, e, e_own);
}

// S = L diag(e) R in row layout.
__device__ inline floatx4 MiddleProduct16(const CompressOperands &o, const float (&e)[4]) {
  float scaled[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) scaled[r] = o.left_rows[r] * e[r];
  return Mfma(o.right_cols, scaled);
}

// Row p of out = L diag(exp(mid)) R, max-normalised with its own FP64 log
// scale (src/inference.cc:749-776).
template <int kItemsPerWave>
__global__ __launch_bounds__(256) void CompressForwardMfma16(DeviceOps ops, Dims d, std::uint32_t offset,
                                                             std::uint32_t count, LinearBuffers buf) {
  const int lane = threadIdx.x % 64;
  const std::uint32_t total = count * d.B;
  const std::uint32_t first = Uniform((blockIdx.x * kWaves + threadIdx.x / 64) * kItemsPerWave);
  if (first >= total) return;
  CompressOperands o[kItemsPerWave];
  double left_scale[kItemsPerWave], mid_scale[kItemsPerWave];
#pragma unroll
  for (int t = 0; t < kItemsPerWave; ++t) {
    const std::uint32_t item = first + t < total ? first + t : first;
    LoadCompression<true>(ops, d, buf, offset, item, lane, o[t]);
    left_scale[t] = RowScale(buf, d, o[t].op.left, o[t].b, lane % 16);
    mid_scale[t] = buf.node_scale[static_cast<std::uint64_t>(o[t].op.middle) * d.B + o[t].b];
  }
#pragma unroll
  for (int t = 0; t < kItemsPerWave; ++t) {
    float e[4], e_own;
    const double w_top = MiddleWeights16(o[t], lane, e, &e_own);
    const floatx4 s = MiddleProduct16(o[t], e);
    float top = fmaxf(fmaxf(s[0], s[1]), fmaxf(s[2], s[3]));
    top = fmaxf(top, __shfl_xor(top, 16, 64));
    top = fmaxf(top, __shfl_xor(top, 32, 64));
    if (first + t >= total) continue;
    const bool live = top > 0.0f && w_top > kNegInf;
    floatx4 x = s;
    const float inverse = RowInverse(top, live, [&](float f) { x *= f; });
    const std::uint64_t out_at = static_cast<std::uint64_t>(o[t].op.out - d.E) * d.B + o[t].b;
    Quad<float> q;
#pragma unroll
    for (int r = 0; r < 4; ++r) q.v[r] = x[r] * inverse;
    StoreQuadStream(static_cast<float *>(buf.paths) + out_at * 256 + (lane % 16) * 16 + 4 * (lane / 16), q);
    // log top to about 6e-8 absolute whatever its magnitude; the sum of
    // scales stays FP64.
    if (lane < 16)
      buf.path_scale[out_at * 16 + lane] =
          live ? left_scale[t] + mid_scale[t] + w_top + LogPositive(top) : kNegInf;
  }
}

// Compression adjoint on MFMA (src/inference.cc:416-464). With S = L' R,
// L' = L diag(e), Gr = adj_out / S (FP64 ratios; 0 where adj_out or S is 0),
// g_p = max_c |Gr[p][c]|, Gh = Gr / g_p, Gamma = max_p g_p, Gt = Gr / Gamma
// (magnitudes: an adjoint seeded by a negative upstream gradient keeps its
// sign through Gh and Gt):
//   adj_L[p][m] = L'[p][m] g_p (Gh R^T)[p][m]
//   adj_R[m][c] = R[m][c] e[m] Gamma (L^T Gt)[m][c]
//   adj_mid[m]  = sum_c adj_R[m][c]
// H^T = R Gh^T takes R and Gh in row layout; Q^T = Gt^T L takes Gt and L in
// column layout, Gt reaching it through one LDS transposition.
constexpr int kPad = 17;  // LDS row stride (floats)

template <int kItemsPerWave>
__global__ __launch_bounds__(256) void CompressReverseMfma16(DeviceOps ops, Dims d, std::uint32_t offset,
                                                             std::uint32_t count, LinearBuffers buf) {
  __shared__ float tiles[kWaves][kItemsPerWave][16 * kPad];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const std::uint32_t total = count * d.B;
  const std::uint32_t first = Uniform((blockIdx.x * kWaves + wave) * kItemsPerWave);
  if (first >= total) return;  // whole wavefront; only wavefront barriers below
  CompressOperands o[kItemsPerWave];
  float right_rows[kItemsPerWave][4], left_cols[kItemsPerWave][4], adj[kItemsPerWave][4];
#pragma unroll
  for (int t = 0; t < kItemsPerWave; ++t) {
    const std::uint32_t item = first + t < total ? first + t : first;
    LoadCompression<false>(ops, d, buf, offset, item, lane, o[t]);
    LoadRows<true>(o[t].right, lane, right_rows[t]);
    LoadColumns<true>(o[t].left, lane, left_cols[t]);
    const std::uint64_t out_at = static_cast<std::uint64_t>(o[t].op.out - d.E) * d.B + o[t].b;
    LoadRows<true>(static_cast<const float *>(buf.path_adj) + out_at * 256, lane, adj[t]);
  }
  float e[kItemsPerWave][4], e_own[kItemsPerWave];
  double g[kItemsPerWave], gamma[kItemsPerWave];
  float gh[kItemsPerWave][4];
#pragma unroll
  for (int t = 0; t < kItemsPerWave; ++t) {
    MiddleWeights16(o[t], lane, e[t], &e_own[t]);
    const floatx4 s = MiddleProduct16(o[t], e[t]);
    double gr[4];
    double top = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      gr[r] = (adj[t][r] != 0.0f && s[r] > 0.0f) ? static_cast<double>(adj[t][r]) / static_cast<double>(s[r]) : 0.0;
      top = fmax(top, fabs(gr[r]));
    }
    top = fmax(top, __shfl_xor(top, 16, 64));
    top = fmax(top, __shfl_xor(top, 32, 64));
    g[t] = top;  // g_p for p = lane % 16
    double big = top;
#pragma unroll
    for (int x = 1; x < 16; x <<= 1) big = fmax(big, __shfl_xor(big, x, 64));
    gamma[t] = big;
    const double inverse_gamma = big > 0.0 ? 1.0 / big : 0.0;
    const double inverse_top = top > 0.0 ? 1.0 / top : 0.0;
    float *tile = tiles[wave][t];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      gh[t][r] = static_cast<float>(gr[r] * inverse_top);
      tile[(lane % 16) * kPad + 4 * (lane / 16) + r] = static_cast<float>(gr[r] * inverse_gamma);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int t = 0; t < kItemsPerWave; ++t) {
    float gt_cols[4];
    const float *tile = tiles[wave][t];
#pragma unroll
    for (int r = 0; r < 4; ++r) gt_cols[r] = tile[(4 * (lane / 16) + r) * kPad + lane % 16];
    const floatx4 h = Mfma(right_rows[t], gh[t]);   // H in row layout: [p = l%16][m = 4(l/16) + r]
    const floatx4 q = Mfma(gt_cols, left_cols[t]);  // Q in row layout: [m = l%16][c = 4(l/16) + r]
    if (first + t >= total) continue;
    const DeviceCompress &op = o[t].op;
    const std::uint32_t b = o[t].b;
    const int col = 4 * (lane / 16);
    Quad<float> adj_left, adj_right;
    double row_total = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      adj_left.v[r] = static_cast<float>(static_cast<double>(o[t].left_rows[r]) * static_cast<double>(e[t][r]) * g[t] *
                                         static_cast<double>(h[r]));
      const double right = static_cast<double>(right_rows[t][r]) * static_cast<double>(e_own[t]) * gamma[t] *
                           static_cast<double>(q[r]);
      adj_right.v[r] = static_cast<float>(right);
      row_total += right;
    }
    StoreAdjointQuad(SlotAdjointTarget<float>(buf, d, op.left, b), (lane % 16) * 16 + col, adj_left);
    StoreAdjointQuad(SlotAdjointTarget<float>(buf, d, op.right, b), (lane % 16) * 16 + col, adj_right);
    row_total += __shfl_xor(row_total, 16, 64);
    row_total += __shfl_xor(row_total, 32, 64);
    if (lane < 16)
      static_cast<float *>(buf.node_adj)[(static_cast<std::uint64_t>(op.middle) * d.B + b) * 16 + lane] =
          static_cast<float>(row_total);
  }
}

// ---------------------------------------------------------------------------
// Grouped compressions, K = 16 (one wavefront per (group, item))
// ---------------------------------------------------------------------------
//
// A group's product X_0 D_1 X_1 ... D_{n-1} X_{n-1} is accumulated left to
// right, T_0 = X_0 and T_j = rownorm(T_{j-1} diag(e_j) X_j), each step being
// the binary compression above with T_{j-1} as its left operand: T stays in
// registers in row layout (the layout of the left operand and of the MFMA
// result), X_j arrives in column layout, and the row scales add up in FP64.
// The reverse pass recomputes T_0 .. T_{n-2} and runs the binary adjoint
// steps from j = n-1 down to 1; the adjoint of T_{j-1} stays in registers as
// the output adjoint of the next step. Adjoints are in marginal form (value
// times gradient), which does not depend on how T's rows are scaled.

struct ChainView {  // wave-uniform
  std::uint32_t n, out, b;
  std::uint32_t slots[kMaxChain], middles[kMaxChain - 1];
};

__device__ inline void LoadChain(const DeviceOps &ops, const Dims &d, std::uint32_t offset, std::uint32_t item,
                                 ChainView &c) {
  item = Uniform(item);
  c.b = item % d.B;
  const DeviceChain &src = ops.chains[offset + item / d.B];
  c.n = Uniform(src.count);
  c.out = Uniform(src.out);
#pragma unroll
  for (std::uint32_t i = 0; i < kMaxChain; ++i) c.slots[i] = Uniform(src.slots[i]);
#pragma unroll
  for (std::uint32_t i = 0; i + 1 < kMaxChain; ++i) c.middles[i] = Uniform(src.middles[i]);
}

__device__ inline float RowMax16(const floatx4 &s) {
  float top = fmaxf(fmaxf(s[0], s[1]), fmaxf(s[2], s[3]));
  top = fmaxf(top, Xor16(top));
  return fmaxf(top, __shfl_xor(top, 32, 64));
}

// Middle node j's value for m = l%16 and the row scale of X_j at m.
__device__ inline void ChainMiddle(const LinearBuffers &buf, const Dims &d, const ChainView &c, int j, int lane,
                                   float *mid_own, double *right_scale_own) {
  const std::uint64_t mid_at = static_cast<std::uint64_t>(c.middles[j - 1]) * d.B + c.b;
  *mid_own = static_cast<const float *>(buf.nodes)[mid_at * 16 + lane % 16];
  *right_scale_own = RowScale(buf, d, c.slots[j], c.b, lane % 16);
}

// `block`: the block's index in the launch (kWaves groups per block).
template <int kMax>
__device__ __forceinline__ void ChainForwardBody(const DeviceOps &ops, const Dims &d, std::uint32_t offset,
                                                 std::uint32_t count, const LinearBuffers &buf, std::uint32_t block) {
  const int lane = threadIdx.x % 64;
  const std::uint32_t item = Uniform(block * kWaves + threadIdx.x / 64);
  if (item >= count * d.B) return;
  ChainView c;
  LoadChain(ops, d, offset, item, c);
  float acc[4], cols[kMax][4], mid_own[kMax];
  double right_scale_own[kMax];
  LoadRows<true>(Slot<float>(buf, d, c.slots[0], c.b), lane, acc);
  double scale = RowScale(buf, d, c.slots[0], c.b, lane % 16);
#pragma unroll
  for (int j = 1; j < kMax; ++j) {
    if (j >= static_cast<int>(c.n)) break;
    LoadColumns<true>(Slot<float>(buf, d, c.slots[j], c.b), lane, cols[j]);
    ChainMiddle(buf, d, c, j, lane, &mid_own[j], &right_scale_own[j]);
  }
#pragma unroll
  for (int j = 1; j < kMax; ++j) {
    if (j >= static_cast<int>(c.n)) break;
    float e[4], e_own;
    const double w_top = StepWeights16(mid_own[j], right_scale_own[j], lane, e, &e_own);
    float scaled[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) scaled[r] = acc[r] * e[r];
    floatx4 s = Mfma(cols[j], scaled);
    const float top = RowMax16(s);
    const bool live = top > 0.0f && w_top > kNegInf;
    const float inverse = RowInverse(top, live, [&](float f) { s *= f; });
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = s[r] * inverse;
    const double mid_scale = buf.node_scale[static_cast<std::uint64_t>(c.middles[j - 1]) * d.B + c.b];
    scale = live ? scale + mid_scale + w_top + LogPositive(top) : kNegInf;
  }
  const std::uint64_t out_at = static_cast<std::uint64_t>(c.out - d.E) * d.B + c.b;
  Quad<float> q;
#pragma unroll
  for (int r = 0; r < 4; ++r) q.v[r] = acc[r];
  StoreQuadStream(static_cast<float *>(buf.paths) + out_at * 256 + (lane % 16) * 16 + 4 * (lane / 16), q);
  if (lane < 16) buf.path_scale[out_at * 16 + lane] = scale;
}

template <int kMax>
__global__ __launch_bounds__(256) void ChainForwardMfma16(DeviceOps ops, Dims d, std::uint32_t offset,
                                                          std::uint32_t count, LinearBuffers buf) {
  ChainForwardBody<kMax>(ops, d, offset, count, buf, blockIdx.x);
}

// Row-layout quad of an adjoint; edge marginals may be caller memory
// without 16-byte alignment (quad_edges == 0), then stored per element.
__device__ inline void StoreAdjointRow(const LinearBuffers &buf, const Dims &d, std::uint32_t slot, std::uint32_t b,
                                       int lane, const Quad<float> &q, bool quad_edges) {
  const AdjointTarget<float> t = SlotAdjointTarget<float>(buf, d, slot, b);
  const std::uint64_t at = (lane % 16) * 16 + 4 * (lane / 16);
  if (slot >= d.E || quad_edges) {
    StoreAdjointQuad(t, at, q);
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) StoreAdjointValue(t, at + r, q.v[r]);
  }
}

// Reverse pass of one group over a balanced bracketing of its product
// (matrix products associate): span [lo, hi) splits at mid = (lo + hi) / 2,
// whose middle node is middles[mid - 1]. Load() then Combine() form every span's
// row-normalised product bottom-up in row and column layout (the operand
// layouts of the adjoint step), so the spans of one depth are independent
// and the critical path is log2(n) adjoint steps. Reverse() runs the binary
// compression adjoint at each split with the span's normalised product P and
// its row normaliser in place of the recomputed S (S = P / normaliser row by
// row), top-down, each split with its own LDS tile for the Gt transposition.
struct Part {
  float row[4], col[4];  // normalised product: row layout, column layout
  double scale;          // FP64 log scale of row l%16 (Load and Combine only)
  float top;             // row l%16 of the product is top * row (1 for a stored slot)
  float e_own;           // split's middle weight e[m], m = l%16
  float e[4];            // split's middle weights e[m], m = 4(l/16) + r (kept when kKeepWeights)
};

struct GroupState {
  const ChainView &c;
  const LinearBuffers &buf;
  const Dims &d;
  int lane;
  float (*tiles)[16 * kPad];  // one per split (index mid - 1)
  bool quad_edges;
};

// kKeepWeights: splits keep their gathered middle weights from Combine for
// Reverse (four registers per split) instead of gathering them again.
template <int kLo, int kHi, bool kKeepWeights, bool kLeaf = (kHi - kLo == 1)> struct Span;

template <int kLo, int kHi, bool kKeepWeights> struct Span<kLo, kHi, kKeepWeights, true> {
  Part part;
  __device__ inline void Load(const GroupState &g) {
    const float *x = Slot<float>(g.buf, g.d, g.c.slots[kLo], g.c.b);
    LoadRows<true>(x, g.lane, part.row);
    LoadColumns<false>(x, g.lane, part.col);
    part.scale = RowScale(g.buf, g.d, g.c.slots[kLo], g.c.b, g.lane % 16);
    part.top = 1.0f;
  }
  __device__ inline void Combine(const GroupState &) {}
  __device__ inline void Reverse(const GroupState &g, const float (&adj)[4]) const {
    Quad<float> q;
#pragma unroll
    for (int r = 0; r < 4; ++r) q.v[r] = adj[r];
    StoreAdjointRow(g.buf, g.d, g.c.slots[kLo], g.c.b, g.lane, q, g.quad_edges);
  }
};

template <int kLo, int kHi, bool kKeepWeights> struct Span<kLo, kHi, kKeepWeights, false> {
  static constexpr int kMid = (kLo + kHi) / 2;
  Span<kLo, kMid, kKeepWeights> left;
  Span<kMid, kHi, kKeepWeights> right;
  Part part;
  float mid_own;      // middle node's value at m = l%16 (Load to Combine)
  double mid_scale;   // its FP64 log scale

  // Every global load of the group (slots, row scales, middle nodes) is
  // issued before the first product: Combine's LDS transpositions order
  // memory, so loads left after one would wait for it.
  __device__ inline void Load(const GroupState &g) {
    left.Load(g);
    right.Load(g);
    const std::uint64_t mid_at = static_cast<std::uint64_t>(g.c.middles[kMid - 1]) * g.d.B + g.c.b;
    mid_own = static_cast<const float *>(g.buf.nodes)[mid_at * 16 + g.lane % 16];
    mid_scale = g.buf.node_scale[mid_at];
  }

  __device__ inline void Combine(const GroupState &g) {
    left.Combine(g);
    right.Combine(g);
    const int lane = g.lane;
    const double w_top = StepWeights16(mid_own, right.part.scale, lane, part.e, &part.e_own);
    float scaled[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) scaled[r] = left.part.row[r] * part.e[r];
    floatx4 s = Mfma(right.part.col, scaled);  // row layout
    const float top = RowMax16(s);
    const bool live = top > 0.0f && w_top > kNegInf;
    const float inverse = RowInverse(top, live, [&](float f) { s *= f; });
    part.top = top;  // unscaled
#pragma unroll
    for (int r = 0; r < 4; ++r) part.row[r] = s[r] * inverse;
    // The column layout is the row layout transposed through this split's
    // LDS tile (Reverse reuses the tile later; one wavefront, in order).
    float *tile = g.tiles[kMid - 1];
#pragma unroll
    for (int r = 0; r < 4; ++r) tile[(lane % 16) * kPad + 4 * (lane / 16) + r] = part.row[r];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int r = 0; r < 4; ++r) part.col[r] = tile[(4 * (lane / 16) + r) * kPad + lane % 16];
    part.scale = live ? left.part.scale + mid_scale + w_top + LogPositive(top) : kNegInf;
  }

  // adj: marginal-form adjoint of this span's product, row layout. The
  // ratios Gr = adj / S (S = top * P, row by row) are formed in FP32 when
  // every one is finite and below 1e30 in magnitude and every row maximum
  // g_p is 0 or at least 2^-100 (so that its reciprocal is finite), else
  // (wave-uniform) in FP64 as in the binary compression adjoint; the row
  // and global scales g_p and Gamma are applied in FP64 when they exceed
  // 1e30.
  __device__ inline void Reverse(const GroupState &g, const float (&adj)[4]) const {
    const int lane = g.lane;
    float e[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) e[r] = kKeepWeights ? part.e[r] : __shfl(part.e_own, 4 * (lane / 16) + r, 64);
    float gr[4];
    bool small = true;
    const float inverse = __builtin_amdgcn_rcpf(part.top);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      gr[r] = (adj[r] != 0.0f && part.row[r] > 0.0f) ? adj[r] * inverse * __builtin_amdgcn_rcpf(part.row[r]) : 0.0f;
      small = small && fabsf(gr[r]) < 1e30f;
    }
    float gh[4], gt[4];
    double top, gamma;
    bool fast = __all(small);
    float t = 0.0f;
    if (fast) {
      t = fmaxf(fmaxf(fabsf(gr[0]), fabsf(gr[1])), fmaxf(fabsf(gr[2]), fabsf(gr[3])));
      t = fmaxf(t, Xor16(t));
      t = fmaxf(t, __shfl_xor(t, 32, 64));
      fast = __all(t == 0.0f || t >= 0x1p-100f);  // then Gamma is 0 or at least 2^-100 as well
    }
    if (fast) {
      const float u = RowLanesMax(t);
      const float inverse_top = t > 0.0f ? __builtin_amdgcn_rcpf(t) : 0.0f;
      const float inverse_gamma = u > 0.0f ? __builtin_amdgcn_rcpf(u) : 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) gh[r] = gr[r] * inverse_top, gt[r] = gr[r] * inverse_gamma;
      top = t, gamma = u;
    } else {
      double wide[4];
      top = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        wide[r] = (adj[r] != 0.0f && part.row[r] > 0.0f)
                      ? static_cast<double>(adj[r]) / (static_cast<double>(part.top) * static_cast<double>(part.row[r]))
                      : 0.0;
        top = fmax(top, fabs(wide[r]));
      }
      top = fmax(top, __shfl_xor(top, 16, 64));
      top = fmax(top, __shfl_xor(top, 32, 64));
      gamma = top;
#pragma unroll
      for (int x = 1; x < 16; x <<= 1) gamma = fmax(gamma, __shfl_xor(gamma, x, 64));
      const double inverse_top = top > 0.0 ? 1.0 / top : 0.0;
      const double inverse_gamma = gamma > 0.0 ? 1.0 / gamma : 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        gh[r] = static_cast<float>(wide[r] * inverse_top), gt[r] = static_cast<float>(wide[r] * inverse_gamma);
    }
    float *tile = g.tiles[kMid - 1];
#pragma unroll
    for (int r = 0; r < 4; ++r) tile[(lane % 16) * kPad + 4 * (lane / 16) + r] = gt[r];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float gt_cols[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) gt_cols[r] = tile[(4 * (lane / 16) + r) * kPad + lane % 16];
    const floatx4 h = Mfma(right.part.row, gh);      // [p = l%16][m = 4(l/16) + r]
    const floatx4 q = Mfma(gt_cols, left.part.col);  // [m = l%16][c = 4(l/16) + r]
    float adj_left[4], adj_right[4], row_total;
    if (__all(top < 1e30 && gamma < 1e30)) {
      const float top_f = static_cast<float>(top), right_weight = part.e_own * static_cast<float>(gamma);
      row_total = 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        adj_left[r] = (left.part.row[r] * e[r]) * (top_f * h[r]);
        adj_right[r] = right.part.row[r] * right_weight * q[r];
        row_total += adj_right[r];
      }
      row_total += Xor16(row_total);
      row_total += __shfl_xor(row_total, 32, 64);
    } else {
      double total = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        adj_left[r] = static_cast<float>(static_cast<double>(left.part.row[r]) * static_cast<double>(e[r]) * top *
                                         static_cast<double>(h[r]));
        const double value = static_cast<double>(right.part.row[r]) * static_cast<double>(part.e_own) * gamma *
                             static_cast<double>(q[r]);
        adj_right[r] = static_cast<float>(value);
        total += value;
      }
      total += __shfl_xor(total, 16, 64);
      total += __shfl_xor(total, 32, 64);
      row_total = static_cast<float>(total);
    }
    if (lane < 16)
      static_cast<float *>(g.buf.node_adj)[(static_cast<std::uint64_t>(g.c.middles[kMid - 1]) * g.d.B + g.c.b) * 16 +
                                           lane] = row_total;
    left.Reverse(g, adj_left);
    right.Reverse(g, adj_right);
  }
};

template <int kN, bool kKeepWeights> __device__ inline void ReverseGroup(const GroupState &g, const float (&adj)[4]) {
  Span<0, kN, kKeepWeights> root;
  root.Load(g);
  root.Combine(g);
  root.Reverse(g, adj);
}

// kMax: the largest group a launch may hold (2, 4 or 8); it bounds the
// registers and LDS tiles of the kernel.
template <int kMax>
__device__ inline void ChainReverseBody(const DeviceOps &ops, const Dims &d, std::uint32_t offset,
                                        std::uint32_t count, const LinearBuffers &buf, int quad_edges,
                                        float (*tiles)[kMax - 1][16 * kPad], std::uint32_t block) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const std::uint32_t item = Uniform(block * kWaves + wave);
  if (item >= count * d.B) return;  // whole wavefront; only wavefront barriers below
  ChainView c;
  LoadChain(ops, d, offset, item, c);
  const GroupState g{c, buf, d, lane, tiles[wave], quad_edges != 0};
  float adj[4];
  const std::uint64_t out_at = static_cast<std::uint64_t>(c.out - d.E) * d.B + c.b;
  LoadRows<true>(static_cast<const float *>(buf.path_adj) + out_at * 256, lane, adj);
  if constexpr (kMax == 2) {
    ReverseGroup<2, true>(g, adj);
    return;
  }
  // Groups of more than 2 slots gather their middle weights again in
  // Reverse (the same shuffles of e_own as in Combine), which keeps the
  // kernel of groups of 3 and 4 at 96 registers: five wavefronts per SIMD.
  switch (c.n) {
    case 2: ReverseGroup<2, false>(g, adj); break;
    case 3: ReverseGroup<3, false>(g, adj); break;
    case 4: ReverseGroup<4, false>(g, adj); break;
    default:
      if constexpr (kMax > 4) {
        switch (c.n) {
          case 5: ReverseGroup<5, false>(g, adj); break;
          case 6: ReverseGroup<6, false>(g, adj); break;
          case 7: ReverseGroup<7, false>(g, adj); break;
          case 8: ReverseGroup<8, false>(g, adj); break;
          default: break;
        }
      }
      break;
  }
}

// Groups of exactly 2 slots.
__global__ __launch_bounds__(256) void ChainReversePair(DeviceOps ops, Dims d, std::uint32_t offset,
                                                        std::uint32_t count, LinearBuffers buf, int quad_edges) {
  __shared__ float tiles[kWaves][1][16 * kPad];
  ChainReverseBody<2>(ops, d, offset, count, buf, quad_edges, tiles, blockIdx.x);
}

// Groups of 3 or 4 slots (the kernel handles 2 as well); five wavefronts
// per SIMD (at most 102 registers per lane).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void ChainReverseSmall(DeviceOps ops, Dims d, std::uint32_t offset,
                                                         std::uint32_t count, LinearBuffers buf, int quad_edges) {
  __shared__ float tiles[kWaves][3][16 * kPad];
  ChainReverseBody<4>(ops, d, offset, count, buf, quad_edges, tiles, blockIdx.x);
}

// Three wavefronts per SIMD (at most 168 registers per lane).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void ChainReverseLarge(
    DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, LinearBuffers buf, int quad_edges) {
  __shared__ float tiles[kWaves][kMaxChain - 1][16 * kPad];
  ChainReverseBody<kMaxChain>(ops, d, offset, count, buf, quad_edges, tiles, blockIdx.x);
}

// ---------------------------------------------------------------------------
// FP32 MFMA compression, K = 32 and 64 (one wavefront per item)
// ---------------------------------------------------------------------------
//
// With NB = K/16 blocks, lane l (i = l%16, g = l/16) holds a K x K matrix X
//   in row layout:    X[16I + i][16J + 4g + r]   (NB*NB 16-byte loads),
//   in column layout: X[16J + 4g + r][16I + i]   (the row layout of X^T).
// Taking step s = (J, r) of the k loop as k = 16J + 4g + r, the MFMA tile
// (Ia, Ib) over two row-layout operands Y, Z accumulates (Y Z^T) with
// D[4g + r][i] = (Y Z^T)[16Ia + 4g + r][16Ib + i]: output row 16Ib + i in
// lane i, columns 16Ia + 4g + r, i.e. the row layout of the product's
// transpose. Hence
//   S^T = R^T (L diag e)^T  from R in column and L in row layout,
//   H^T = R Gh^T            from R and Gh in row layout,
//   Q^T = Gt^T L            from Gt and L in column layout,
// and S, H, Q come out in row layout: rows reduce over 4 lanes, stores are
// 16-byte quads. Gh/Gt pass through one LDS tile per wavefront.

// Middle weights: lane l forms e[m] for m = l % K (*e_own), given the scale
// of row m of the right slot, and gathers those of m = 16J + 4g + r into
// e[J][r]; returns W (-inf: no live state).
template <int K>
__device__ inline double MiddleWeightsWide(const LinearBuffers &buf, const Dims &d, const DeviceCompress &op,
                                           std::uint32_t b, int lane, double right_scale, float (&e)[K / 16][4],
                                           float *e_own) {
  constexpr int NB = K / 16;
  const int g = lane / 16;
  const std::uint64_t mid_at = static_cast<std::uint64_t>(op.middle) * d.B + b;
  const int m = lane % K;
  const double w = static_cast<double>(static_cast<const float *>(buf.nodes)[mid_at * K + m]) + right_scale;
  double top = w;
#pragma unroll
  for (int x = 1; x < K; x <<= 1) top = fmax(top, __shfl_xor(top, x, 64));
  *e_own = top > kNegInf ? ExpNonPositive(static_cast<float>(w - top)) : 0.0f;
#pragma unroll
  for (int J = 0; J < NB; ++J)
#pragma unroll
    for (int r = 0; r < 4; ++r) e[J][r] = __shfl(*e_own, 16 * J + 4 * g + r, 64);
  return top;
}

// kStream: the kernel's last read of these values (non-temporal loads).
template <int K, bool kStream = false>
__device__ inline void LoadRowBand(const float *X, int I, int i, int g, float (&out)[K / 16][4]) {
#pragma unroll
  for (int J = 0; J < K / 16; ++J) {
    const float *at = X + (16 * I + i) * K + 16 * J + 4 * g;
    const Quad<float> q = kStream ? LoadQuadStream(at) : LoadQuad(at);
#pragma unroll
    for (int r = 0; r < 4; ++r) out[J][r] = q.v[r];
  }
}

template <int K, bool kStream = false>
__device__ inline void LoadColumnBand(const float *X, int I, int i, int g, float (&out)[K / 16][4]) {
#pragma unroll
  for (int J = 0; J < K / 16; ++J)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float *at = X + (16 * J + 4 * g + r) * K + 16 * I + i;
      out[J][r] = kStream ? __builtin_nontemporal_load(at) : *at;
    }
}

// The row-shifted exponential of a log edge value, as edge staging forms it:
// exp(x - m) in a row of maximum m, 0 in a row of -inf.
__device__ inline float ShiftedExp(float x, float m) { return m > NegInf<float>() ? ExpNonPositive(x - m) : 0.0f; }

// A compression input that is an edge read as log potentials in place.
__device__ inline bool LogEdgeInput(const LinearBuffers &buf, const Dims &d, std::uint32_t slot) {
  return buf.compress_log_edges != 0 && slot < d.E;
}

__device__ inline const float *CompressInput(const LinearBuffers &buf, const Dims &d, std::uint32_t slot,
                                             std::uint32_t b) {
  return LogEdgeInput(buf, d, slot)
             ? static_cast<const float *>(buf.log_edges) + static_cast<std::uint64_t>(slot) * d.K * d.K
             : Slot<float>(buf, d, slot, b);
}

// The op of a wavefront's item with its fields in scalar registers.
__device__ inline DeviceCompress UniformOp(const DeviceOps &ops, std::uint32_t index) {
  const DeviceCompress op = ops.compresses[index];
  DeviceCompress u{};
  u.left = Uniform(op.left), u.middle = Uniform(op.middle), u.right = Uniform(op.right), u.out = Uniform(op.out);
  return u;
}

template <int K>
__device__ inline floatx4 TileProduct(const float (&y)[K / 16][4], const float (&z)[K / 16][4]) {
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int J = 0; J < K / 16; ++J)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(y[J][r], z[J][r], acc, 0, 0, 0);
  return acc;
}

constexpr int kWideWaves = 2;  // wavefronts (items) per block

template <int K>
__global__ __launch_bounds__(64 * kWideWaves) void CompressForwardMfmaWide(DeviceOps ops, Dims d, std::uint32_t offset,
                                                                           std::uint32_t count, LinearBuffers buf) {
  constexpr int NB = K / 16;
  __shared__ float right_shift[kWideWaves][K];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, i = lane % 16, g = lane / 16;
  const std::uint32_t item = Uniform(blockIdx.x * kWideWaves + wave);
  if (item >= count * d.B) return;  // whole wavefront; only wavefront barriers below
  const std::uint32_t b = item % d.B;
  const DeviceCompress op = UniformOp(ops, offset + item / d.B);
  const bool log_left = LogEdgeInput(buf, d, op.left), log_right = LogEdgeInput(buf, d, op.right);
  const float *L = CompressInput(buf, d, op.left, b);
  const float *R = CompressInput(buf, d, op.right, b);
  float rc[NB][NB][4];  // R column layout: rc[Ic] is the band Ic of R^T
#pragma unroll
  for (int Ic = 0; Ic < NB; ++Ic) LoadColumnBand<K, true>(R, Ic, i, g, rc[Ic]);
  double right_scale;  // of row lane % K of R
  if (log_right) {
    // Row 16J + 4g + r of R lies in the 16 lanes of group g, NB values each.
#pragma unroll
    for (int J = 0; J < NB; ++J)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float top = rc[0][J][r];
#pragma unroll
        for (int Ic = 1; Ic < NB; ++Ic) top = fmaxf(top, rc[Ic][J][r]);
#pragma unroll
        for (int x = 1; x < 16; x <<= 1) top = fmaxf(top, __shfl_xor(top, x, 64));
#pragma unroll
        for (int Ic = 0; Ic < NB; ++Ic) rc[Ic][J][r] = ShiftedExp(rc[Ic][J][r], top);
        if (i == 0) right_shift[wave][16 * J + 4 * g + r] = top;
      }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float top = right_shift[wave][lane % K];
    right_scale = top > NegInf<float>() ? static_cast<double>(top) : kNegInf;
    // Edges are shared by the batch: the op's items at other batch indices
    // store the same values here, and nothing reads them in this pass.
    if (lane < K) buf.log_edge_shift[static_cast<std::uint64_t>(op.right) * K + lane] = right_scale;
  } else {
    right_scale = RowScale(buf, d, op.right, b, lane % K);
  }
  float e[NB][4], e_own;
  const double w_top = MiddleWeightsWide<K>(buf, d, op, b, lane, right_scale, e, &e_own);
  const std::uint64_t mid_at = static_cast<std::uint64_t>(op.middle) * d.B + b;
  const double mid_scale = buf.node_scale[mid_at];
  const std::uint64_t out_at = static_cast<std::uint64_t>(op.out - d.E) * d.B + b;
  float *out = static_cast<float *>(buf.paths) + out_at * K * K;
#pragma unroll 1
  for (int Ip = 0; Ip < NB; ++Ip) {
    float lb[NB][4];
    LoadRowBand<K, true>(L, Ip, i, g, lb);
    double left_scale;
    if (log_left) {  // row 16Ip + i lies in lanes i, i + 16, i + 32, i + 48
      float top = lb[0][0];
#pragma unroll
      for (int J = 0; J < NB; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) top = fmaxf(top, lb[J][r]);
      top = fmaxf(top, __shfl_xor(top, 16, 64));
      top = fmaxf(top, __shfl_xor(top, 32, 64));
#pragma unroll
      for (int J = 0; J < NB; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) lb[J][r] = ShiftedExp(lb[J][r], top);
      left_scale = top > NegInf<float>() ? static_cast<double>(top) : kNegInf;
      if (g == 0) buf.log_edge_shift[static_cast<std::uint64_t>(op.left) * K + 16 * Ip + i] = left_scale;
    } else {
      left_scale = RowScale(buf, d, op.left, b, 16 * Ip + i);
    }
#pragma unroll
    for (int J = 0; J < NB; ++J)
#pragma unroll
      for (int r = 0; r < 4; ++r) lb[J][r] *= e[J][r];
    floatx4 s[NB];
    float top = 0.0f;
#pragma unroll
    for (int Ic = 0; Ic < NB; ++Ic) {
      s[Ic] = TileProduct<K>(rc[Ic], lb);  // S[16Ip + i][16Ic + 4g + r]
      top = fmaxf(top, fmaxf(fmaxf(s[Ic][0], s[Ic][1]), fmaxf(s[Ic][2], s[Ic][3])));
    }
    top = fmaxf(top, __shfl_xor(top, 16, 64));
    top = fmaxf(top, __shfl_xor(top, 32, 64));
    const bool live = top > 0.0f && w_top > kNegInf;
    const float inverse = RowInverse(top, live, [&](float f) {
#pragma unroll
      for (int Ic = 0; Ic < NB; ++Ic) s[Ic] *= f;
    });
#pragma unroll
    for (int Ic = 0; Ic < NB; ++Ic) {
      Quad<float> q;
#pragma unroll
      for (int r = 0; r < 4; ++r) q.v[r] = s[Ic][r] * inverse;
      StoreQuadStream(out + (16 * Ip + i) * K + 16 * Ic + 4 * g, q);
    }
    if (g == 0)
      buf.path_scale[out_at * K + 16 * Ip + i] =
          live ? left_scale + mid_scale + w_top + LogPositive(top) : kNegInf;
  }
}

template <int K>
__global__ __launch_bounds__(64 * kWideWaves) void CompressReverseMfmaWide(DeviceOps ops, Dims d, std::uint32_t offset,
                                                                           std::uint32_t count, LinearBuffers buf) {
  constexpr int NB = K / 16;
  constexpr int kStride = K + 1;  // LDS row stride (floats)
  __shared__ float tiles[kWideWaves][K * kStride];
  __shared__ double row_ratio[kWideWaves][K];
  __shared__ float left_shift[kWideWaves][K];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, i = lane % 16, g = lane / 16;
  const std::uint32_t item = Uniform(blockIdx.x * kWideWaves + wave);
  if (item >= count * d.B) return;  // whole wavefront; only wavefront barriers below
  float *G = tiles[wave];
  double *gs = row_ratio[wave];
  float *ls = left_shift[wave];
  const std::uint32_t b = item % d.B;
  const DeviceCompress op = UniformOp(ops, offset + item / d.B);
  // Log edge inputs are shifted by the row maxima the forward pass stored
  // as their row scales, which gives back the forward's values bit for bit.
  const bool log_left = LogEdgeInput(buf, d, op.left), log_right = LogEdgeInput(buf, d, op.right);
  const float *L = CompressInput(buf, d, op.left, b);
  const float *R = CompressInput(buf, d, op.right, b);
  const std::uint64_t out_at = static_cast<std::uint64_t>(op.out - d.E) * d.B + b;
  const float *adj_out = static_cast<const float *>(buf.path_adj) + out_at * K * K;
  float e[NB][4], e_own;
  const double w_top =
      MiddleWeightsWide<K>(buf, d, op, b, lane, RowScale(buf, d, op.right, b, lane % K), e, &e_own);
  const double mid_scale = buf.node_scale[static_cast<std::uint64_t>(op.middle) * d.B + b];

  // S is not recomputed: the forward pass stored P = S / top_p row by row
  // as the compression's output slot and log top_p inside its row scale
  // (left scale + middle scale + W + log top_p, summed in that order).
  // Gr' = adj/P = Gr top_p and g_p' = max_c |Gr'| give Gh = Gr/g_p =
  // Gr'/g_p' (-> LDS) and g_p = g_p' / top_p (FP64); Gamma = max_p g_p. A
  // band's ratios are formed in FP32 (hardware reciprocal) when every one
  // is finite and below 1e30 in magnitude and every g_p' is 0 or at least
  // 2^-100 (a finite reciprocal), else (wave-uniform) in FP64.
  const float *P = static_cast<const float *>(buf.paths) + out_at * K * K;
  double gamma = 0.0;
#pragma unroll 1
  for (int Ip = 0; Ip < NB; ++Ip) {
    float pb[NB][4], ab[NB][4];
    LoadRowBand<K, true>(P, Ip, i, g, pb);
    LoadRowBand<K, true>(adj_out, Ip, i, g, ab);
    const int row = 16 * Ip + i;
    const double path_scale = buf.path_scale[out_at * K + row];
    const double left_scale = RowScale(buf, d, op.left, b, row);
    float gr[NB][4];
    float top_f = 0.0f;
    bool small = true;
#pragma unroll
    for (int Ic = 0; Ic < NB; ++Ic)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        gr[Ic][r] = (ab[Ic][r] != 0.0f && pb[Ic][r] > 0.0f) ? ab[Ic][r] * __builtin_amdgcn_rcpf(pb[Ic][r]) : 0.0f;
        small = small && fabsf(gr[Ic][r]) < 1e30f;
        top_f = fmaxf(top_f, fabsf(gr[Ic][r]));
      }
    double top;  // g_p'
    bool fast = __all(small);
    if (fast) {
      top_f = fmaxf(top_f, __shfl_xor(top_f, 16, 64));
      top_f = fmaxf(top_f, __shfl_xor(top_f, 32, 64));
      fast = __all(top_f == 0.0f || top_f >= 0x1p-100f);
    }
    if (fast) {
      const float inverse = top_f > 0.0f ? __builtin_amdgcn_rcpf(top_f) : 0.0f;
#pragma unroll
      for (int Ic = 0; Ic < NB; ++Ic)
#pragma unroll
        for (int r = 0; r < 4; ++r) G[row * kStride + 16 * Ic + 4 * g + r] = gr[Ic][r] * inverse;
      top = top_f;
    } else {
      double wide[NB][4];
      top = 0.0;
#pragma unroll
      for (int Ic = 0; Ic < NB; ++Ic)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          wide[Ic][r] = (ab[Ic][r] != 0.0f && pb[Ic][r] > 0.0f)
                            ? static_cast<double>(ab[Ic][r]) / static_cast<double>(pb[Ic][r])
                            : 0.0;
          top = fmax(top, fabs(wide[Ic][r]));
        }
      top = fmax(top, __shfl_xor(top, 16, 64));
      top = fmax(top, __shfl_xor(top, 32, 64));
      const double inverse = top > 0.0 ? 1.0 / top : 0.0;
#pragma unroll
      for (int Ic = 0; Ic < NB; ++Ic)
#pragma unroll
        for (int r = 0; r < 4; ++r) G[row * kStride + 16 * Ic + 4 * g + r] = static_cast<float>(wide[Ic][r] * inverse);
    }
    // A row with a nonzero ratio is live in the forward pass: its scales
    // are finite.
    const double g_row = top > 0.0 ? top / exp(path_scale - ((left_scale + mid_scale) + w_top)) : 0.0;
    gamma = fmax(gamma, g_row);
    if (g == 0) {
      gs[row] = g_row;
      ls[row] = static_cast<float>(left_scale);  // exact: an FP32 maximum or -inf
    }
  }
#pragma unroll
  for (int x = 1; x < 16; x <<= 1) gamma = fmax(gamma, __shfl_xor(gamma, x, 64));
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  float rr[NB][NB][4];  // R row layout: rr[Im] is the band Im of R
#pragma unroll
  for (int Im = 0; Im < NB; ++Im) {
    LoadRowBand<K, true>(R, Im, i, g, rr[Im]);
    if (log_right) {
      const float top = static_cast<float>(RowScale(buf, d, op.right, b, 16 * Im + i));
#pragma unroll
      for (int J = 0; J < NB; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) rr[Im][J][r] = ShiftedExp(rr[Im][J][r], top);
    }
  }

  // adj_L[p][m] = L'[p][m] g_p H[p][m], H^T = R Gh^T.
  const AdjointTarget<float> adj_left = SlotAdjointTarget<float>(buf, d, op.left, b);
#pragma unroll 1
  for (int Ip = 0; Ip < NB; ++Ip) {
    float gb[NB][4], lb[NB][4];
#pragma unroll
    for (int J = 0; J < NB; ++J)
#pragma unroll
      for (int r = 0; r < 4; ++r) gb[J][r] = G[(16 * Ip + i) * kStride + 16 * J + 4 * g + r];
    LoadRowBand<K>(L, Ip, i, g, lb);
    if (log_left) {
      const float top = ls[16 * Ip + i];
#pragma unroll
      for (int J = 0; J < NB; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) lb[J][r] = ShiftedExp(lb[J][r], top);
    }
    const double g_row = gs[16 * Ip + i];
    // FP32 products when every row's g_p is below 1e30 (|H| <= K, L and e
    // at most 1), else FP64.
    const bool narrow = __all(g_row < 1e30);
    const float g_row_f = static_cast<float>(g_row);
#pragma unroll
    for (int Im = 0; Im < NB; ++Im) {
      const floatx4 h = TileProduct<K>(rr[Im], gb);  // H[16Ip + i][16Im + 4g + r]
      Quad<float> q;
      if (narrow) {
#pragma unroll
        for (int r = 0; r < 4; ++r) q.v[r] = (lb[Im][r] * e[Im][r]) * (g_row_f * h[r]);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          q.v[r] = static_cast<float>(static_cast<double>(lb[Im][r]) * static_cast<double>(e[Im][r]) * g_row *
                                      static_cast<double>(h[r]));
      }
      StoreAdjointQuad(adj_left, (16 * Ip + i) * K + 16 * Im + 4 * g, q);
    }
  }

  // adj_R[m][c] = R[m][c] e[m] Gamma Q[m][c], Q^T = Gt^T L, Gt = Gr / Gamma;
  // adj_mid[m] = sum_c adj_R[m][c].
  const double inverse_gamma = gamma > 0.0 ? 1.0 / gamma : 0.0;
  float scale[NB][4];  // g_p / Gamma for p = 16J + 4g + r
#pragma unroll
  for (int J = 0; J < NB; ++J)
#pragma unroll
    for (int r = 0; r < 4; ++r) scale[J][r] = static_cast<float>(gs[16 * J + 4 * g + r] * inverse_gamma);
  const AdjointTarget<float> adj_right = SlotAdjointTarget<float>(buf, d, op.right, b);
  const std::uint64_t mid_at = static_cast<std::uint64_t>(op.middle) * d.B + b;
  // Unrolled: rr is indexed by Im and must stay in registers.
#pragma unroll
  for (int Im = 0; Im < NB; ++Im) {
    float lc[NB][4];
    LoadColumnBand<K, true>(L, Im, i, g, lc);
    if (log_left) {  // lc[J][r] is in row 16J + 4g + r of L
#pragma unroll
      for (int J = 0; J < NB; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) lc[J][r] = ShiftedExp(lc[J][r], ls[16 * J + 4 * g + r]);
    }
    // e[m] for m = 16Im + i (the row this lane holds of Q and adj_R).
    const float e_row = __shfl(e_own, 16 * Im + i, 64);
    double row_total = 0.0;
#pragma unroll
    for (int Ic = 0; Ic < NB; ++Ic) {
      float gt[NB][4];
#pragma unroll
      for (int J = 0; J < NB; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) gt[J][r] = G[(16 * J + 4 * g + r) * kStride + 16 * Ic + i] * scale[J][r];
      const floatx4 q = TileProduct<K>(gt, lc);  // Q[16Im + i][16Ic + 4g + r]
      Quad<float> out;
      if (gamma < 1e30) {  // wave-uniform; |Q| <= K, R and e at most 1
        const float weight = e_row * static_cast<float>(gamma);
        float quad_total = 0.0f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          out.v[r] = rr[Im][Ic][r] * weight * q[r];
          quad_total += out.v[r];
        }
        row_total += static_cast<double>(quad_total);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double v = static_cast<double>(rr[Im][Ic][r]) * static_cast<double>(e_row) * gamma *
                           static_cast<double>(q[r]);
          out.v[r] = static_cast<float>(v);
          row_total += v;
        }
      }
      StoreAdjointQuad(adj_right, (16 * Im + i) * K + 16 * Ic + 4 * g, out);
    }
    row_total += __shfl_xor(row_total, 16, 64);
    row_total += __shfl_xor(row_total, 32, 64);
    if (g == 0) static_cast<float *>(buf.node_adj)[mid_at * K + 16 * Im + i] = static_cast<float>(row_total);
  }
}

// ---------------------------------------------------------------------------
// Vector transposition [X][Y][K] -> [Y][X][K], K in {4, 8, 16, 32, 64}
// ---------------------------------------------------------------------------

// A team of K/4 lanes moves one K-vector with 16-byte (float) or 32-byte
// (double) accesses; consecutive teams take consecutive input vectors, so
// reads are contiguous and, for X = 1 or Y = 1 (a forest, or one tree), so
// are writes. kMode 0 copies; kMode 1 takes logs of linear potentials and
// kMode 2 log potentials, both then shifted by the vector's maximum, which
// goes to `scale` in FP64 (-inf for a vector without support); kMode 3 turns
// log potentials into exp(v - max) with the same FP64 scale (edge rows).
// Lane l of a team moves input vector v to output vector `to`.
template <typename In, typename Out, int kMode, int K>
__device__ inline void TransposeVector(const In *in, Out *out, std::uint64_t v, std::uint64_t to, int l,
                                       double *scale) {
  constexpr int kLanes = K / 4;
  const Quad<In> q = LoadQuad(in + v * K + l * 4);
  // Logs are taken in the wider of the two element types.
  using W = typename std::conditional<sizeof(In) >= sizeof(Out), In, Out>::type;
  Quad<Out> o;
  if constexpr (kMode == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) o.v[i] = static_cast<Out>(q.v[i]);
  } else {
    W w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = kMode == 1 ? LogOf(static_cast<W>(q.v[i])) : static_cast<W>(q.v[i]);
    W top = Max(Max(w[0], w[1]), Max(w[2], w[3]));
#pragma unroll
    for (int s = kLanes / 2; s > 0; s >>= 1) top = Max(top, __shfl_xor(top, s, kLanes));
    const bool live = top > NegInf<W>();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if constexpr (kMode == 3)
        o.v[i] = live ? ExpNonPositive(static_cast<Out>(w[i] - top)) : Out(0);
      else
        o.v[i] = live ? static_cast<Out>(w[i] - top) : NegInf<Out>();
    }
    if (l == 0) scale[to] = live ? static_cast<double>(top) : kNegInf;
  }
  StoreQuad(out + to * K + l * 4, o);
}

template <typename In, typename Out, int kMode, int K>
__global__ __launch_bounds__(256) void TransposeTeam(const In *in, Out *out, std::uint32_t X, std::uint32_t Y,
                                                     double *scale) {
  constexpr int kLanes = K / 4;
  // Grid-stride: the stride is a whole number of teams, so a team's lanes
  // stay together.
  for (std::uint64_t tid = GridThread(); tid < static_cast<std::uint64_t>(X) * Y * kLanes; tid += GridStride()) {
    const std::uint64_t v = tid / kLanes;
    const std::uint64_t x = v / Y, y = v % Y;
    TransposeVector<In, Out, kMode, K>(in, out, v, y * X + x, static_cast<int>(tid % kLanes), scale);
  }
}

// The same for the vectors y = ids[i] (i < count) only, for every x.
template <typename In, typename Out, int kMode, int K>
__global__ __launch_bounds__(256) void ListedTransposeTeam(const In *in, Out *out, std::uint32_t X, std::uint32_t Y,
                                                           const std::uint32_t *ids, std::uint32_t count,
                                                           double *scale) {
  constexpr int kLanes = K / 4;
  for (std::uint64_t tid = GridThread(); tid < static_cast<std::uint64_t>(count) * X * kLanes; tid += GridStride()) {
    const std::uint64_t team = tid / kLanes;
    const std::uint64_t i = team / X, x = team % X;
    const std::uint64_t y = ids[i];
    TransposeVector<In, Out, kMode, K>(in, out, x * Y + y, y * X + x, static_cast<int>(tid % kLanes), scale);
  }
}

constexpr unsigned kBlock = 256;

bool Aligned16(const void *p) { return (reinterpret_cast<std::uintptr_t>(p) & 15u) == 0; }

unsigned Blocks(std::uint64_t threads) { return static_cast<unsigned>((threads + kBlock - 1) / kBlock); }

hipStream_t StreamOf(const Launch &l) { return static_cast<hipStream_t>(l.stream); }

unsigned BlocksOf(std::uint64_t threads, unsigned block) { return static_cast<unsigned>((threads + block - 1) / block); }

template <typename T, int K> void LaunchRakeTeam(const Launch &l, bool reverse, std::uint32_t offset, std::uint32_t count,
                                                 const LinearBuffers &buf) {
  const unsigned block = l.tuning.rake_block;
  const unsigned grid = BlocksOf(static_cast<std::uint64_t>(count) * l.dims.B * Team<K>::kLanes, block);
  if (reverse && buf.raw_leaves != nullptr)
    RakeReverseTeam<T, K, false, true><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
  else if (reverse && buf.log_edges != nullptr)
    RakeReverseTeam<T, K, true><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
  else if (reverse)
    RakeReverseTeam<T, K, false><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
  else
    RakeForwardTeam<T, K><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
}

// ---------------------------------------------------------------------------
// Persistent passes (K = 16, FP32, batch 1)
// ---------------------------------------------------------------------------

__global__ void ClearWords(std::uint32_t *words, std::uint32_t count) {
  for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) words[i] = 0;
}

// Thread 0 takes the next task and waits until its group has completed the
// task's earlier steps; the block reads the index from LDS.
//
// Agent scope spans the device's XCDs, whose L2 caches are separate, so an
// agent-scope acquire invalidates the XCD's L2 and a release writes it back.
// The wait therefore polls with relaxed loads (no cache maintenance) and
// takes one acquire when the wait is over; a task ends with one release.
// The wait is bounded in wall-clock time (q.wait_ticks of the constant
// wall clock), after which the launch sets its failure word and stops
// waiting, so a broken dependency cannot hang the device.
__device__ inline std::uint32_t TakeTask(const PersistentQueue &q, std::uint32_t *next) {
  if (threadIdx.x == 0) {
    const std::uint32_t i = atomicAdd(q.counters, 1u);
    *next = i;
    if (i < q.num_tasks) {
      const DeviceTask t = q.tasks[i];
      const std::uint32_t *progress = q.counters + 1 + q.steps + t.group;
      if (__hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < t.step) {
        const std::uint64_t start = wall_clock64();
        while (__hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < t.step) {
          if (__hip_atomic_load(q.failure, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
          if (wall_clock64() - start > q.wait_ticks) {
            atomicOr(q.failure, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(8);
        }
      }
      // One invalidation (the CU's L1, shared by the block's wavefronts,
      // and the XCD's L2) before the block reads what other tasks wrote.
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
  }
  __syncthreads();
  return *next;
}

// After a task: every wavefront waits for its stores to reach the L2, then
// one release makes them visible device-wide before the task is counted;
// the last task of its (group, step) advances the group's progress.
__device__ inline void FinishTask(const PersistentQueue &q, const DeviceTask &t) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const std::uint32_t at = q.need_offset[t.group] + t.step;
    if (atomicAdd(q.counters + 1 + at, 1u) + 1 == q.need[at])
      __hip_atomic_store(q.counters + 1 + q.steps + t.group, t.step + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The fused forward program (rake chunks, partial sums, compression groups
// of up to 4 slots), each task run as blocks of the level kernels would.
template <bool kLogEdges>
__global__ __launch_bounds__(256) void PersistentForward16(GatherOps g, DeviceOps ops, Dims d, LinearBuffers buf,
                                                           PersistentQueue q) {
  __shared__ std::uint32_t next;
  for (;;) {
    const std::uint32_t i = TakeTask(q, &next);
    if (i >= q.num_tasks) return;
    const DeviceTask t = q.tasks[i];
    for (std::uint32_t j = 0; j < t.blocks; ++j) {
      if (t.kind == kTaskGather)
        RakeGatherTeamBody<float, 16, kLogEdges>(g, d, t.offset, t.count, buf, j * 256 + threadIdx.x);
      else if (t.kind == kTaskSegments)
        GatherSegmentsTeamBody<float, 16>(g, d, t.offset, t.count, buf, j * 256 + threadIdx.x);
      else
        ChainForwardBody<4>(ops, d, t.offset, t.count, buf, j);
    }
    FinishTask(q, t);
  }
}

// The reverse program: rake adjoints and compression groups of 2 and of 3
// or 4 slots.
template <bool kLogEdges>
__global__ __launch_bounds__(256) void PersistentReverse16(DeviceOps ops, Dims d, LinearBuffers buf,
                                                           PersistentQueue q, int quad_edges) {
  __shared__ std::uint32_t next;
  __shared__ float tiles[kWaves][3][16 * kPad];
  for (;;) {
    const std::uint32_t i = TakeTask(q, &next);
    if (i >= q.num_tasks) return;
    const DeviceTask t = q.tasks[i];
    for (std::uint32_t j = 0; j < t.blocks; ++j) {
      if (t.kind == kTaskRakeReverse)
        RakeReverseTeamBody<float, 16, kLogEdges>(ops, d, t.offset, t.count, buf, j * 256 + threadIdx.x);
      else if (t.kind == kTaskChainReversePair)
        ChainReverseBody<2>(ops, d, t.offset, t.count, buf, quad_edges,
                            reinterpret_cast<float (*)[1][16 * kPad]>(&tiles[0][0][0]), j);
      else
        ChainReverseBody<4>(ops, d, t.offset, t.count, buf, quad_edges, tiles, j);
    }
    FinishTask(q, t);
  }
}

template <typename F> std::uint32_t ResidentGrid(F kernel) {
  int device = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&device) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess)
    return 0;
  return static_cast<std::uint32_t>(cus) * static_cast<std::uint32_t>(per_cu);
}

}  // namespace

bool FusedForwardSupported(std::uint32_t states) {
  return states == 8 || states == 16 || states == 32 || states == 64;
}

void LaunchGatherChunks(const Launch &l, const GatherOps &g, std::uint32_t offset, std::uint32_t count,
                        const LinearBuffers &buf) {
  if (count == 0) return;
  if (static_cast<std::uint64_t>(count) * l.dims.B * 64 >= kMaxLaunchItems) {
    InOpChunks(offset, count, l.dims.B,
               [&](std::uint32_t first, std::uint32_t ops) { LaunchGatherChunks(l, g, first, ops, buf); });
    return;
  }
  const std::uint32_t K = l.dims.K;
  const unsigned block = l.tuning.rake_block;
  const unsigned grid = BlocksOf(static_cast<std::uint64_t>(count) * l.dims.B * (K >= 32 ? 64 : 16), block);
  const hipStream_t s = StreamOf(l);
#define THMM_GATHER(T, KK)                                                                  \
  do {                                                                                      \
    if (buf.raw_leaves != nullptr)                                                          \
      RakeGatherTeam<T, KK, false, true><<<grid, block, 0, s>>>(g, l.dims, offset, count, buf); \
    else if (buf.log_edges != nullptr)                                                      \
      RakeGatherTeam<T, KK, true><<<grid, block, 0, s>>>(g, l.dims, offset, count, buf);    \
    else                                                                                    \
      RakeGatherTeam<T, KK, false><<<grid, block, 0, s>>>(g, l.dims, offset, count, buf);   \
  } while (0)
  if (l.real == Real::kF32) {
    if (K == 8) THMM_GATHER(float, 8);
    if (K == 16) THMM_GATHER(float, 16);
    if (K == 32) THMM_GATHER(float, 32);
    if (K == 64) THMM_GATHER(float, 64);
  } else {
    if (K == 8) THMM_GATHER(double, 8);
    if (K == 16) THMM_GATHER(double, 16);
    if (K == 32) THMM_GATHER(double, 32);
    if (K == 64) THMM_GATHER(double, 64);
  }
#undef THMM_GATHER
  RecordLaunchError(hipGetLastError(), "fused rake level launch");
}

void LaunchGatherSegments(const Launch &l, const GatherOps &g, std::uint32_t offset, std::uint32_t count,
                          const LinearBuffers &buf) {
  if (count == 0) return;
  if (static_cast<std::uint64_t>(count) * l.dims.B * 64 >= kMaxLaunchItems) {
    InOpChunks(offset, count, l.dims.B,
               [&](std::uint32_t first, std::uint32_t ops) { LaunchGatherSegments(l, g, first, ops, buf); });
    return;
  }
  const std::uint32_t K = l.dims.K;
  const unsigned grid = Blocks(static_cast<std::uint64_t>(count) * l.dims.B * K);
  const hipStream_t s = StreamOf(l);
#define THMM_SEGMENTS(T, KK)                                                                  \
  do {                                                                                        \
    if (buf.raw_leaves != nullptr)                                                            \
      GatherSegmentsTeam<T, KK, true><<<grid, kBlock, 0, s>>>(g, l.dims, offset, count, buf); \
    else                                                                                      \
      GatherSegmentsTeam<T, KK><<<grid, kBlock, 0, s>>>(g, l.dims, offset, count, buf);       \
  } while (0)
  if (l.real == Real::kF32) {
    if (K == 8) THMM_SEGMENTS(float, 8);
    if (K == 16) THMM_SEGMENTS(float, 16);
    if (K == 32) THMM_SEGMENTS(float, 32);
    if (K == 64) THMM_SEGMENTS(float, 64);
  } else {
    if (K == 8) THMM_SEGMENTS(double, 8);
    if (K == 16) THMM_SEGMENTS(double, 16);
    if (K == 32) THMM_SEGMENTS(double, 32);
    if (K == 64) THMM_SEGMENTS(double, 64);
  }
#undef THMM_SEGMENTS
  RecordLaunchError(hipGetLastError(), "partial absorption launch");
}

std::uint64_t WallClockTicks(double seconds) {
  int device = 0, khz = 0;
  if (hipGetDevice(&device) != hipSuccess ||
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0)
    khz = 100000;  // the 100 MHz clock of gfx9 parts
  return static_cast<std::uint64_t>(seconds * 1e3 * khz);
}

std::uint32_t PersistentGrid(bool reverse) {
  const std::uint32_t a = reverse ? ResidentGrid(PersistentReverse16<true>) : ResidentGrid(PersistentForward16<true>);
  const std::uint32_t b = reverse ? ResidentGrid(PersistentReverse16<false>) : ResidentGrid(PersistentForward16<false>);
  return a < b ? a : b;
}

void LaunchPersistentPass(const Launch &l, bool reverse, const GatherOps &g, const PersistentQueue &q,
                          const LinearBuffers &buf) {
  if (q.num_tasks == 0) return;
  const hipStream_t s = StreamOf(l);
  const bool log_edges = buf.log_edges != nullptr;
  const std::uint32_t words = 1 + q.steps + q.groups;
  ClearWords<<<(words + kBlock - 1) / kBlock < 1024 ? (words + kBlock - 1) / kBlock : 1024, kBlock, 0, s>>>(
      q.counters, words);
  const unsigned grid = q.grid == 0 ? 1u : (q.grid < q.num_tasks ? q.grid : q.num_tasks);
  if (reverse) {
    const int quad_edges = Aligned16(buf.edge_out);
    if (log_edges)
      PersistentReverse16<true><<<grid, kBlock, 0, s>>>(l.ops, l.dims, buf, q, quad_edges);
    else
      PersistentReverse16<false><<<grid, kBlock, 0, s>>>(l.ops, l.dims, buf, q, quad_edges);
  } else {
    if (log_edges)
      PersistentForward16<true><<<grid, kBlock, 0, s>>>(g, l.ops, l.dims, buf, q);
    else
      PersistentForward16<false><<<grid, kBlock, 0, s>>>(g, l.ops, l.dims, buf, q);
  }
  RecordLaunchError(hipGetLastError(), "persistent pass launch");
}

bool LaunchFastLevel(const Launch &l, bool reverse, std::uint32_t primitive, std::uint32_t offset,
                     std::uint32_t count, const LinearBuffers &buf) {
  const std::uint32_t K = l.dims.K;
  if (primitive == kChainPrimitive || primitive == kSmallChainPrimitive || primitive == kPairChainPrimitive) {
    // Built only for K = 16 FP32 workspaces; there is no generic fallback
    // (the inner compressions' slots are not materialised).
    if (K != 16 || l.real != Real::kF32 || l.ops.chains == nullptr) {
      RecordLaunchError(hipErrorInvalidValue, "grouped compression launch (K = 16 FP32 only)");
      return true;
    }
    const unsigned grid = static_cast<unsigned>((static_cast<std::uint64_t>(count) * l.dims.B + kWaves - 1) / kWaves);
    const hipStream_t s = StreamOf(l);
    if (reverse && primitive == kChainPrimitive)
      ChainReverseLarge<<<grid, kBlock, 0, s>>>(l.ops, l.dims, offset, count, buf, Aligned16(buf.edge_out));
    else if (reverse && primitive == kSmallChainPrimitive)
      ChainReverseSmall<<<grid, kBlock, 0, s>>>(l.ops, l.dims, offset, count, buf, Aligned16(buf.edge_out));
    else if (reverse)
      ChainReversePair<<<grid, kBlock, 0, s>>>(l.ops, l.dims, offset, count, buf, Aligned16(buf.edge_out));
    else
      ChainForwardMfma16<kMaxChain><<<grid, kBlock, 0, s>>>(l.ops, l.dims, offset, count, buf);
    RecordLaunchError(hipGetLastError(), "grouped compression launch");
    return true;
  }
  // 32-bit item indexing: the grid must address every (op, item, lane).
  // Caller-owned edge inputs and outputs must allow 16-byte accesses.
  if (!Aligned16(buf.edges) || !Aligned16(buf.edge_out) || !Aligned16(buf.node_adj)) return false;
  if (static_cast<std::uint64_t>(count) * l.dims.B * 64 >= (1ull << 32)) return false;
  if (primitive == 0 && (K == 8 || K == 16 || K == 32 || K == 64)) {
    if (l.real == Real::kF32) {
      if (K == 8) LaunchRakeTeam<float, 8>(l, reverse, offset, count, buf);
      if (K == 16) LaunchRakeTeam<float, 16>(l, reverse, offset, count, buf);
      if (K == 32) LaunchRakeTeam<float, 32>(l, reverse, offset, count, buf);
      if (K == 64) LaunchRakeTeam<float, 64>(l, reverse, offset, count, buf);
    } else {
      if (K == 8) LaunchRakeTeam<double, 8>(l, reverse, offset, count, buf);
      if (K == 16) LaunchRakeTeam<double, 16>(l, reverse, offset, count, buf);
      if (K == 32) LaunchRakeTeam<double, 32>(l, reverse, offset, count, buf);
      if (K == 64) LaunchRakeTeam<double, 64>(l, reverse, offset, count, buf);
    }
    RecordLaunchError(hipGetLastError(), "rake level launch");
    return true;
  }
  if (primitive == 3 && (K == 32 || K == 64) && l.real == Real::kF32) {
    const unsigned grid =
        static_cast<unsigned>((static_cast<std::uint64_t>(count) * l.dims.B + kWideWaves - 1) / kWideWaves);
    const hipStream_t s = StreamOf(l);
    if (K == 32) {
      if (reverse)
        CompressReverseMfmaWide<32><<<grid, 64 * kWideWaves, 0, s>>>(l.ops, l.dims, offset, count, buf);
      else
        CompressForwardMfmaWide<32><<<grid, 64 * kWideWaves, 0, s>>>(l.ops, l.dims, offset, count, buf);
    } else {
      if (reverse)
        CompressReverseMfmaWide<64><<<grid, 64 * kWideWaves, 0, s>>>(l.ops, l.dims, offset, count, buf);
      else
        CompressForwardMfmaWide<64><<<grid, 64 * kWideWaves, 0, s>>>(l.ops, l.dims, offset, count, buf);
    }
    RecordLaunchError(hipGetLastError(), "MFMA compression launch");
    return true;
  }
  if (primitive == 3 && K == 16 && l.real == Real::kF32) {
    const int items = reverse ? l.tuning.compress_items_reverse : l.tuning.compress_items_forward;
    const std::uint64_t per_block = static_cast<std::uint64_t>(kWaves) * items;
    const unsigned grid = static_cast<unsigned>((static_cast<std::uint64_t>(count) * l.dims.B + per_block - 1) / per_block);
    const hipStream_t s = StreamOf(l);
    if (reverse && items == 1) CompressReverseMfma16<1><<<grid, kBlock, 0, s>>>(l.ops, l.dims, offset, count, buf);
    if (reverse && items != 1) CompressReverseMfma16<2><<<grid, kBlock, 0, s>>>(l.ops, l.dims, offset, count, buf);
    if (!reverse && items == 1) CompressForwardMfma16<1><<<grid, kBlock, 0, s>>>(l.ops, l.dims, offset, count, buf);
    if (!reverse && items != 1) CompressForwardMfma16<2><<<grid, kBlock, 0, s>>>(l.ops, l.dims, offset, count, buf);
    RecordLaunchError(hipGetLastError(), "MFMA compression launch");
    return true;
  }
  return false;
}

namespace {

template <typename In, typename Out, int kMode>
bool TransposeTeamFor(const Launch &l, const void *in, void *out, std::uint32_t X, std::uint32_t Y, double *scale,
                      const NodeList *only = nullptr) {
  const std::uint64_t vectors = only != nullptr ? static_cast<std::uint64_t>(only->count) * X
                                                : static_cast<std::uint64_t>(X) * Y;
  if (vectors == 0) return true;
  const unsigned grid = GridStrideBlocks(vectors * (l.dims.K / 4), kBlock);
  const auto *i = static_cast<const In *>(in);
  auto *o = static_cast<Out *>(out);
  const hipStream_t s = StreamOf(l);
#define THMM_TRANSPOSE(KK)                                                                                        \
  do {                                                                                                            \
    if (only != nullptr)                                                                                          \
      ListedTransposeTeam<In, Out, kMode, KK><<<grid, kBlock, 0, s>>>(i, o, X, Y, only->ids, only->count, scale); \
    else                                                                                                          \
      TransposeTeam<In, Out, kMode, KK><<<grid, kBlock, 0, s>>>(i, o, X, Y, scale);                               \
  } while (0)
  switch (l.dims.K) {
    case 4: THMM_TRANSPOSE(4); break;
    case 8: THMM_TRANSPOSE(8); break;
    case 16: THMM_TRANSPOSE(16); break;
    case 32: THMM_TRANSPOSE(32); break;
    case 64: THMM_TRANSPOSE(64); break;
    default: return false;
  }
#undef THMM_TRANSPOSE
  RecordLaunchError(hipGetLastError(), "vector transpose launch");
  return true;
}

}  // namespace

namespace {

// Mode-3 staging (exp of each row shifted by its maximum, with that
// maximum as the row's FP64 scale) of the listed edges only, in place:
// row r of the launch is row r % K of edge ids[r / K].
template <typename T, int K>
__global__ __launch_bounds__(256) void StageLogRowsTeam(const T *log_edges, const std::uint32_t *ids,
                                                        std::uint64_t rows, T *edges, double *scale) {
  constexpr int kLanes = K / 4;
  for (std::uint64_t tid = GridThread(); tid < rows * kLanes; tid += GridStride()) {  // whole teams per stride
    const std::uint64_t r = tid / kLanes;
    const int l = static_cast<int>(tid % kLanes);
    const std::uint64_t v = static_cast<std::uint64_t>(ids[r / K]) * K + r % K;
    const Quad<T> q = LoadQuad(log_edges + v * K + l * 4);
    T top = Max(Max(q.v[0], q.v[1]), Max(q.v[2], q.v[3]));
#pragma unroll
    for (int s = kLanes / 2; s > 0; s >>= 1) top = Max(top, __shfl_xor(top, s, kLanes));
    const bool live = top > NegInf<T>();
    Quad<T> o;
#pragma unroll
    for (int i = 0; i < 4; ++i) o.v[i] = live ? ExpNonPositive(static_cast<T>(q.v[i] - top)) : T(0);
    StoreQuad(edges + v * K + l * 4, o);
    if (l == 0) scale[v] = live ? static_cast<double>(top) : kNegInf;
  }
}

}  // namespace

bool LaunchStageLogEdges(const Launch &l, const void *log_edges, const std::uint32_t *ids, std::uint32_t count,
                         void *edges, double *edge_scale) {
  const std::uint32_t K = l.dims.K;
  if (K != 8 && K != 16 && K != 32 && K != 64) return false;
  if (count == 0) return true;
  const std::uint64_t rows = static_cast<std::uint64_t>(count) * K;
  const unsigned grid = GridStrideBlocks(rows * (K / 4), kBlock);
  const hipStream_t s = StreamOf(l);
#define THMM_STAGE(T, KK)                                                                                        \
  StageLogRowsTeam<T, KK><<<grid, kBlock, 0, s>>>(static_cast<const T *>(log_edges), ids, rows,                  \
                                                  static_cast<T *>(edges), edge_scale)
  if (l.real == Real::kF32) {
    if (K == 8) THMM_STAGE(float, 8);
    if (K == 16) THMM_STAGE(float, 16);
    if (K == 32) THMM_STAGE(float, 32);
    if (K == 64) THMM_STAGE(float, 64);
  } else {
    if (K == 8) THMM_STAGE(double, 8);
    if (K == 16) THMM_STAGE(double, 16);
    if (K == 32) THMM_STAGE(double, 32);
    if (K == 64) THMM_STAGE(double, 64);
  }
#undef THMM_STAGE
  RecordLaunchError(hipGetLastError(), "log edge staging launch");
  return true;
}

bool FastTransposeFits(std::uint32_t K, std::uint32_t X, std::uint32_t Y) {
  if (K % 4 != 0 || K > 64 || (K & (K - 1)) != 0) return false;
  // Below 8 states a vector is under 32 bytes: scattered writes only pay when
  // one side of the transposition is trivial.
  return K >= 8 || X == 1 || Y == 1;
}

bool LaunchFastTranspose(const Launch &l, int mode, Real in_real, Real out_real, const void *in, void *out,
                         std::uint32_t X, std::uint32_t Y, double *scale, const NodeList *only) {
  const std::uint32_t K = l.dims.K;
  if (!FastTransposeFits(K, X, Y)) return false;
  if (!Aligned16(in) || !Aligned16(out)) return false;
  if (static_cast<std::uint64_t>(X) * Y == 0) return true;
  const bool f_in = in_real == Real::kF32, f_out = out_real == Real::kF32;
  if (only != nullptr && mode != 1) return false;
  switch (mode) {
    case 0:
      if (f_in != f_out) return false;
      return f_in ? TransposeTeamFor<float, float, 0>(l, in, out, X, Y, scale)
                  : TransposeTeamFor<double, double, 0>(l, in, out, X, Y, scale);
    case 1:
      if (f_in && f_out) return TransposeTeamFor<float, float, 1>(l, in, out, X, Y, scale, only);
      if (f_in) return TransposeTeamFor<float, double, 1>(l, in, out, X, Y, scale, only);
      if (f_out) return TransposeTeamFor<double, float, 1>(l, in, out, X, Y, scale, only);
      return TransposeTeamFor<double, double, 1>(l, in, out, X, Y, scale, only);
    case 2:
      if (f_in != f_out) return false;
      return f_in ? TransposeTeamFor<float, float, 2>(l, in, out, X, Y, scale)
                  : TransposeTeamFor<double, double, 2>(l, in, out, X, Y, scale);
    case 3:
      if (f_in != f_out) return false;
      return f_in ? TransposeTeamFor<float, float, 3>(l, in, out, X, Y, scale)
                  : TransposeTeamFor<double, double, 3>(l, in, out, X, Y, scale);
  }
  return false;
}

}  // namespace thmm
