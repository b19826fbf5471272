// gfx950 kernels of the batched tree-HMM contraction.
//
// Work decomposition: one "item" is one op of a plan level applied to one
// batch item. A group of G lanes (G = next power of two >= K, at most 64)
// owns an item, lane s owning state s (the parent state of a rake, the
// output row of a compression). Items are numbered op-major, batch-minor, so
// the lanes of a wavefront share the op and read a batch-shared edge tile
// once while their node/branch vectors are contiguous in [node][batch][state]
// storage.
//
// Number representation (element type T = float for the FP32 form, double
// for the plain FP64 form and for sampling):
//  * node and branch vectors are max-shifted logs (largest entry 0) plus an
//    FP64 log scale, so products of messages of any dynamic range are sums;
//  * composed path matrices are linear with every row divided by its
//    maximum, plus one FP64 log scale per row ("per-row max shift"); edge
//    slots hold the caller's linear potentials (row scale 0) or, for log
//    inputs, row-max-shifted exponentials with their row scales;
//  * adjoints are linear (posterior masses in [0, 1]).
// A rake is then a linear mat-vec on exp(shifted log), a compression a
// linear matrix product after folding the right factor's row scales into the
// middle vector, and every normaliser is a logarithm added to an FP64 scale.
// This is the scaled sum-product of src/inference.cc:697-836 with the
// normalisation moved into the log domain; the reverse pass is the analytic
// adjoint of src/inference.cc:416-518 written with ratios of normalised
// terms (term / output instead of exp(term - output)), the ratios taken in
// FP64. Max-product follows src/inference.cc:564-695 in FP64; sampling
// follows the inverse-CDF rule of src/inference.cc:128-146 in FP64.
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "device_common.h"
#include "launch.h"
#include "log_f32.h"

namespace thmm {
namespace {
thread_local std::string g_launch_error;
}  // namespace

void RecordLaunchError(hipError_t status, const char *what) {
  if (status != hipSuccess && g_launch_error.empty())
    g_launch_error = std::string(what) + ": " + hipGetErrorString(status);
}

namespace {

using namespace dev;

void Check(hipError_t status, const char *what) { RecordLaunchError(status, what); }

struct Lane {
  std::uint64_t item;
  std::uint32_t s;
};

template <int G> __device__ inline Lane LaneOf() {
  const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  return {tid / G, static_cast<std::uint32_t>(tid % G)};
}

// ------------------------------------------------------------------------
// Scaled forward
// ------------------------------------------------------------------------

// log v for v > 0, else -inf (also for NaN), in the element type.
__device__ inline float ElementLog(float v) { return v > 0.0f ? logf(v) : NegInf<float>(); }

// Entry s of the max-shifted log leaf vector of categorical node n, item b
// (lane s of a G-lane group; NegInf for s >= K) and the vector's FP64 scale,
// formed as InitCategorical forms them: the product in T, its log in FP64
// (or in FP32 with c.element_logs), the maximum over the group.
template <typename T, int G>
__device__ inline T CategoricalLeaf(const CategoricalLeaves &c, const Dims &d, std::uint32_t n, std::uint32_t b,
                                    std::uint32_t s, double *scale) {
  const bool rooted = c.is_root[n] != 0;
  const std::uint32_t o = c.index[n];
  const std::uint32_t category = o == 0xFFFFFFFFu ? 0 : c.obs[static_cast<std::uint64_t>(o) * d.B + b];
  const T *root = static_cast<const T *>(c.root);
  double x = 0.0;
  if (s < d.K) {
    T value;
    if (o == 0xFFFFFFFFu) {
      value = rooted ? root[s] : T(1);
    } else {
      const T e = category < c.categories
                      ? static_cast<const T *>(c.emissions)[static_cast<std::uint64_t>(category) * d.K + s]
                      : T(0);
      value = (rooted ? root[s] : T(1)) * e;
    }
    x = static_cast<double>(value);
  }
  if constexpr (std::is_same<T, float>::value) {
    if (c.element_logs) {
      const float l = s < d.K ? ElementLog(static_cast<float>(x)) : NegInf<float>();
      const float top = GroupMax<G>(l);
      const bool live = top > NegInf<float>();
      *scale = live ? static_cast<double>(top) : kNegInf;
      return live ? l - top : NegInf<float>();
    }
  }
  const double l = x > 0.0 ? log(x) : kNegInf;
  const double top = GroupMax<G>(l);
  *scale = top;
  return top > kNegInf && x > 0.0 ? static_cast<T>(l - top) : NegInf<T>();
}

// exp of the leaf vector of a rake: entry c of exp(v) is read from the node
// state, or for a categorical leaf (`own` = this lane's exp(v[s]), the same
// flag on every lane of the group) taken from lane c of the group.
template <typename T, int G>
__device__ inline T LeafExp(bool categorical, T own, const T *v, std::uint32_t c) {
  return categorical ? __shfl(own, static_cast<int>(c), G) : Exp(v[c]);
}

// branch[p] = log sum_c path[p][c] exp(leaf[c]) (src/inference.cc:704-723)
template <typename T, int G>
__global__ void RakeForward(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, LinearBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const std::uint32_t r = offset + static_cast<std::uint32_t>(lane.item / d.B);
  const DeviceRake op = ops.rakes[r];
  const std::uint64_t leaf_at = static_cast<std::uint64_t>(op.leaf) * d.B + b;
  const T *v = static_cast<const T *>(buf.nodes) + leaf_at * d.K;
  const bool categorical = r < buf.cat.rakes;
  T own = T(0);
  double leaf_scale = kNegInf;
  if (categorical) {
    const T x = CategoricalLeaf<T, G>(buf.cat, d, op.leaf, b, lane.s, &leaf_scale);
    own = Exp(x);
  } else {
    leaf_scale = buf.node_scale[leaf_at];
  }
  // Lanes s >= K run the loop on row K - 1 (shuffles need the whole group)
  // and drop the result.
  const std::uint32_t s = lane.s < d.K ? lane.s : d.K - 1;
  const T *row = Slot<T>(buf, d, op.slot, b) + static_cast<std::uint64_t>(s) * d.K;
  T acc = T(0);
  for (std::uint32_t c = 0; c < d.K; ++c) acc += row[c] * LeafExp<T, G>(categorical, own, v, c);
  double message = kNegInf;
  if (lane.s < d.K && acc > T(0)) message = RowScale(buf, d, op.slot, b, lane.s) + log(static_cast<double>(acc));
  const double top = GroupMax<G>(message);
  const std::uint64_t at = static_cast<std::uint64_t>(op.branch) * d.B + b;
  const bool live = top > kNegInf;
  if (lane.s < d.K) static_cast<T *>(buf.branches)[at * d.K + lane.s] = live ? static_cast<T>(message - top) : NegInf<T>();
  if (lane.s == 0) buf.branch_scale[at] = live ? leaf_scale + top : kNegInf;
}

// Sum of two shifted log vectors, re-shifted (src/inference.cc:725-747).
template <typename T, int G>
__device__ inline void AddLogVectors(T *dst, double *dst_scale, const T *src, double src_scale, std::uint32_t s,
                                     std::uint32_t K) {
  const T x = s < K ? dst[s] + src[s] : NegInf<T>();
  const T top = GroupMax<G>(x);
  const bool live = top > NegInf<T>();
  if (s < K) dst[s] = live ? x - top : NegInf<T>();
  if (s == 0) *dst_scale = live ? *dst_scale + src_scale + static_cast<double>(top) : kNegInf;
}

template <typename T, int G>
__global__ void CombineForward(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, LinearBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const DeviceCombine op = ops.combines[offset + lane.item / d.B];
  const std::uint64_t dst = static_cast<std::uint64_t>(op.destination) * d.B + b;
  const std::uint64_t src = static_cast<std::uint64_t>(op.source) * d.B + b;
  T *branches = static_cast<T *>(buf.branches);
  AddLogVectors<T, G>(branches + dst * d.K, buf.branch_scale + dst, branches + src * d.K, buf.branch_scale[src],
                      lane.s, d.K);
}

template <typename T, int G>
__global__ void AbsorbForward(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, LinearBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const DeviceAbsorb op = ops.absorbs[offset + lane.item / d.B];
  const std::uint64_t node = static_cast<std::uint64_t>(op.parent) * d.B + b;
  const std::uint64_t branch = static_cast<std::uint64_t>(op.branch) * d.B + b;
  AddLogVectors<T, G>(static_cast<T *>(buf.nodes) + node * d.K, buf.node_scale + node,
                      static_cast<const T *>(buf.branches) + branch * d.K, buf.branch_scale[branch], lane.s, d.K);
}

// CombineForward (kAbsorb false) and AbsorbForward (true) at K = 4 with one
// lane per (op, item): AddLogVectors' arithmetic, the vectors moved as
// quads. The maximum over the four sums is fmax's, which is the same in any
// order.
template <typename T, bool kAbsorb>
__global__ __launch_bounds__(256) void AddForwardK4(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count,
                                                    LinearBuffers buf) {
  constexpr int K = 4;
  const std::uint64_t item = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (item >= static_cast<std::uint64_t>(count) * d.B) return;
  const std::uint32_t b = static_cast<std::uint32_t>(item % d.B);
  const std::uint32_t i = offset + static_cast<std::uint32_t>(item / d.B);
  T *dst;
  double *dst_scale;
  std::uint64_t src;
  if constexpr (kAbsorb) {
    const DeviceAbsorb op = ops.absorbs[i];
    const std::uint64_t node = static_cast<std::uint64_t>(op.parent) * d.B + b;
    dst = static_cast<T *>(buf.nodes) + node * K;
    dst_scale = buf.node_scale + node;
    src = static_cast<std::uint64_t>(op.branch) * d.B + b;
  } else {
    const DeviceCombine op = ops.combines[i];
    const std::uint64_t to = static_cast<std::uint64_t>(op.destination) * d.B + b;
    dst = static_cast<T *>(buf.branches) + to * K;
    dst_scale = buf.branch_scale + to;
    src = static_cast<std::uint64_t>(op.source) * d.B + b;
  }
  const Quad<T> a = LoadQuad(dst);
  const Quad<T> c = LoadQuad(static_cast<const T *>(buf.branches) + src * K);
  T x[K];
#pragma unroll
  for (int s = 0; s < K; ++s) x[s] = a.v[s] + c.v[s];
  const T top = Max(Max(x[0], x[1]), Max(x[2], x[3]));
  const bool live = top > NegInf<T>();
  Quad<T> q;
#pragma unroll
  for (int s = 0; s < K; ++s) q.v[s] = live ? x[s] - top : NegInf<T>();
  StoreQuad(dst, q);
  *dst_scale = live ? *dst_scale + buf.branch_scale[src] + static_cast<double>(top) : kNegInf;
}

// Middle weights of a compression: exp(mid[m] + r_m - W) with r_m the right
// factor's row scales folded in and W their maximum (returned; -inf when no
// middle state is supported).
template <typename T, int G>
__device__ inline double MiddleWeights(const LinearBuffers &buf, const Dims &d, const DeviceCompress &op,
                                       std::uint32_t b, const T *mid, T *e) {
  double top = kNegInf;
  for (std::uint32_t m = 0; m < d.K; ++m) top = fmax(top, static_cast<double>(mid[m]) + RowScale(buf, d, op.right, b, m));
#pragma unroll
  for (int m = 0; m < G; ++m) {
    if (m < static_cast<int>(d.K)) {
      const double w = static_cast<double>(mid[m]) + RowScale(buf, d, op.right, b, m);
      e[m] = top > kNegInf ? static_cast<T>(exp(w - top)) : T(0);
    }
  }
  return top;
}

// Row p of out = L diag(exp(mid)) R, each row max-normalised with its own
// FP64 log scale (src/inference.cc:749-776).
template <typename T, int G>
__global__ void CompressForward(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, LinearBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B || lane.s >= d.K) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const DeviceCompress op = ops.compresses[offset + lane.item / d.B];
  const std::uint64_t p = lane.s;
  const T *L = Slot<T>(buf, d, op.left, b) + p * d.K;
  const T *R = Slot<T>(buf, d, op.right, b);
  const std::uint64_t mid_at = static_cast<std::uint64_t>(op.middle) * d.B + b;
  T e[G];
  const double w_top = MiddleWeights<T, G>(buf, d, op, b, static_cast<const T *>(buf.nodes) + mid_at * d.K, e);
  T acc[G];
#pragma unroll
  for (int c = 0; c < G; ++c) acc[c] = T(0);
  for (std::uint32_t m = 0; m < d.K; ++m) {
    const T lm = L[m] * e[m];
    const T *row = R + static_cast<std::uint64_t>(m) * d.K;
#pragma unroll
    for (int c = 0; c < G; ++c)
      if (c < static_cast<int>(d.K)) acc[c] += lm * row[c];
  }
  T top = T(0);
#pragma unroll
  for (int c = 0; c < G; ++c)
    if (c < static_cast<int>(d.K)) top = Max(top, acc[c]);
  const std::uint64_t out_at = static_cast<std::uint64_t>(op.out - d.E) * d.B + b;
  T *out = static_cast<T *>(buf.paths) + out_at * d.K * d.K + p * d.K;
  const bool live = top > T(0) && w_top > kNegInf;
  const T inverse = live ? T(1) / top : T(0);
#pragma unroll
  for (int c = 0; c < G; ++c)
    if (c < static_cast<int>(d.K)) out[c] = acc[c] * inverse;
  buf.path_scale[out_at * d.K + p] =
      live ? RowScale(buf, d, op.left, b, lane.s) + buf.node_scale[mid_at] + w_top + log(static_cast<double>(top))
           : kNegInf;
}

// Team form of CompressForward at K = 4, 8, 16 and 32: K*K/4 lanes per
// (op, item), each owning a quad of four consecutive columns of one output
// row. The middle weights, L diag(e) and R are staged in LDS with quad
// loads; rows are reduced across the K/4 lanes that hold them. The
// arithmetic is CompressForward's, in the same order and element type, so
// the results are bitwise those of the generic kernel.
template <typename T, int K>
__global__ __launch_bounds__(256) void CompressForwardTeam(DeviceOps ops, Dims d, std::uint32_t offset,
                                                           std::uint32_t count, LinearBuffers buf) {
  constexpr int kQ = K / 4, kLanes = K * kQ, kItems = 256 / kLanes;
  __shared__ double weights[kItems][K];  // mid[m] + r_m
  __shared__ T left[kItems][K * K];      // L[p][m] e[m]
  __shared__ T right[kItems][K * K];
  const int t = threadIdx.x / kLanes, l = threadIdx.x % kLanes;
  const std::uint64_t item = static_cast<std::uint64_t>(blockIdx.x) * kItems + t;
  const bool active = item < static_cast<std::uint64_t>(count) * d.B;
  DeviceCompress op{};
  std::uint32_t b = 0;
  std::uint64_t mid_at = 0;
  if (active) {
    b = static_cast<std::uint32_t>(item % d.B);
    op = ops.compresses[offset + item / d.B];
    mid_at = static_cast<std::uint64_t>(op.middle) * d.B + b;
    if (l < K)
      weights[t][l] = static_cast<double>(static_cast<const T *>(buf.nodes)[mid_at * K + l]) +
                      RowScale(buf, d, op.right, b, l);
  }
  __syncthreads();
  double top_w = kNegInf;
  if (active) {
    for (int m = 0; m < K; ++m) top_w = fmax(top_w, weights[t][m]);
    const Quad<T> lq = LoadSlotQuad<T>(buf, d, op.left, b, 4 * l);
    const Quad<T> rq = LoadSlotQuad<T>(buf, d, op.right, b, 4 * l);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = (4 * l + i) % K;
      const T e = top_w > kNegInf ? static_cast<T>(exp(weights[t][m] - top_w)) : T(0);
      left[t][4 * l + i] = lq.v[i] * e;
      right[t][4 * l + i] = rq.v[i];
    }
  }
  __syncthreads();
  if (!active) return;
  const int p = l / kQ, c0 = (l % kQ) * 4;
  T acc[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int m = 0; m < K; ++m) {
    const T lm = left[t][p * K + m];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] += lm * right[t][m * K + c0 + c];
  }
  T top = Max(Max(T(0), Max(acc[0], acc[1])), Max(acc[2], acc[3]));
#pragma unroll
  for (int o = 1; o < kQ; o <<= 1) top = Max(top, __shfl_xor(top, o, kQ));
  const bool live = top > T(0) && top_w > kNegInf;
  const T inverse = live ? T(1) / top : T(0);
  const std::uint64_t out_at = static_cast<std::uint64_t>(op.out - d.E) * d.B + b;
  Quad<T> q;
#pragma unroll
  for (int c = 0; c < 4; ++c) q.v[c] = acc[c] * inverse;
  StoreQuad(static_cast<T *>(buf.paths) + out_at * K * K + p * K + c0, q);
  if (c0 == 0)
    buf.path_scale[out_at * K + p] =
        live ? RowScale(buf, d, op.left, b, p) + buf.node_scale[mid_at] + top_w + log(static_cast<double>(top))
             : kNegInf;
}

// Team form of CompressReverse at K = 4, 8 and 16: K*K/4 lanes per (op,
// item), lane (p, quad q) owning entries 4q..4q+3 of row p. L diag(e), L, R
// and the ratios Gr are staged in LDS; then lane (p, q) forms row p of S and
// Gr on its quad, adj_L[p][4q..] and, as lane (m = p, q), adj_R[m][4q..].
// The FP64 form is CompressReverse's arithmetic, in the same order and
// element types (FP64 ratios and sums), so its results are bitwise those of
// the generic kernel; the middle adjoint sums each row of adj_R
// sequentially in LDS.
// FP32 (T = float): the ratios are taken with the hardware reciprocal and
// the sums kept in FP32 when, for every lane of the block, each S entry is
// 0 or at least 2^-100 and each path adjoint below 2^26 in magnitude (so a
// ratio stays below 2^126) and every FP32 result is finite; otherwise
// (block-uniform) the block takes the FP64 form. The FP32 values share the
// FP64 arrays' LDS.
template <typename T, int K>
__global__ __launch_bounds__(256) void CompressReverseTeam(DeviceOps ops, Dims d, std::uint32_t offset,
                                                           std::uint32_t count, LinearBuffers buf) {
  constexpr int kQ = K / 4, kLanes = K * kQ, kItems = 256 / kLanes;
  constexpr bool kFloat = std::is_same<T, float>::value;
  __shared__ double weights[kItems][K];
  __shared__ T left[kItems][K * K];   // L[p][m] e[m]
  __shared__ T lraw[kItems][K * K];   // L
  __shared__ T right[kItems][K * K];  // R
  __shared__ double gr[kItems][K * K];
  __shared__ double adjr[kItems][K * K];
  const int t = threadIdx.x / kLanes, l = threadIdx.x % kLanes;
  const int p = l / kQ, c0 = (l % kQ) * 4;
  const std::uint64_t item = static_cast<std::uint64_t>(blockIdx.x) * kItems + t;
  const bool active = item < static_cast<std::uint64_t>(count) * d.B;
  float *gr32 = reinterpret_cast<float *>(&gr[t][0]);
  float *adjr32 = reinterpret_cast<float *>(&adjr[t][0]);
  DeviceCompress op{};
  std::uint32_t b = 0;
  std::uint64_t mid_at = 0, out_at = 0;
  if (active) {
    b = static_cast<std::uint32_t>(item % d.B);
    op = ops.compresses[offset + item / d.B];
    mid_at = static_cast<std::uint64_t>(op.middle) * d.B + b;
    out_at = static_cast<std::uint64_t>(op.out - d.E) * d.B + b;
    if (l < K)
      weights[t][l] = static_cast<double>(static_cast<const T *>(buf.nodes)[mid_at * K + l]) +
                      RowScale(buf, d, op.right, b, l);
  }
  __syncthreads();
  double top_w = kNegInf;
  T e_row = T(0);  // e[p]
  Quad<T> a{};
  if (active) {
    for (int m = 0; m < K; ++m) top_w = fmax(top_w, weights[t][m]);
    const Quad<T> lq = LoadSlotQuad<T>(buf, d, op.left, b, 4 * l);
    const Quad<T> rq = LoadSlotQuad<T>(buf, d, op.right, b, 4 * l);
    a = LoadQuad(static_cast<const T *>(buf.path_adj) + out_at * K * K + 4 * l);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = (4 * l + i) % K;
      const T e = top_w > kNegInf ? static_cast<T>(exp(weights[t][m] - top_w)) : T(0);
      left[t][4 * l + i] = lq.v[i] * e;
      lraw[t][4 * l + i] = lq.v[i];
      right[t][4 * l + i] = rq.v[i];
    }
    e_row = top_w > kNegInf ? static_cast<T>(exp(weights[t][p] - top_w)) : T(0);
  }
  __syncthreads();
  T acc[4] = {T(0), T(0), T(0), T(0)};
  bool fast = kFloat;
  if constexpr (kFloat) {
    bool ok = true;
    if (active) {
#pragma unroll
      for (int m = 0; m < K; ++m) {
        const T lm = left[t][p * K + m];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] += lm * right[t][m * K + c0 + c];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        ok = ok && !(acc[c] > 0.0f && acc[c] < 0x1p-100f) && fabsf(a.v[c]) < 0x1p26f;
        gr32[p * K + c0 + c] =
            (a.v[c] != 0.0f && acc[c] > 0.0f) ? a.v[c] * __builtin_amdgcn_rcpf(acc[c]) : 0.0f;
      }
    }
    fast = __syncthreads_and(ok) != 0;
    Quad<T> ql, qr;
    float row_total = 0.0f;
    bool finite = true;
    if (fast && active) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = c0 + i;
        float h = 0.0f;
#pragma unroll
        for (int c = 0; c < K; ++c) h += gr32[p * K + c] * right[t][m * K + c];
        ql.v[i] = left[t][p * K + m] * h;
      }
      const int m = p;
      float r[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int pp = 0; pp < K; ++pp) {
        const float lp = lraw[t][pp * K + m];
#pragma unroll
        for (int c = 0; c < 4; ++c) r[c] += lp * gr32[pp * K + c0 + c];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) qr.v[c] = right[t][m * K + c0 + c] * e_row * r[c];
#pragma unroll
      for (int c = 0; c < 4; ++c) finite = finite && isfinite(ql.v[c]) && isfinite(qr.v[c]);
    }
    if (fast) fast = __syncthreads_and(finite) != 0;  // block-uniform
    if (fast) {
      if (active) {
        StoreAdjointQuad<T>(SlotAdjointTarget<T>(buf, d, op.left, b), static_cast<std::uint64_t>(p) * K + c0, ql);
        StoreAdjointQuad<T>(SlotAdjointTarget<T>(buf, d, op.right, b), static_cast<std::uint64_t>(p) * K + c0, qr);
#pragma unroll
        for (int c = 0; c < 4; ++c) adjr32[p * K + c0 + c] = qr.v[c];
      }
      __syncthreads();
      if (active && c0 == 0) {
#pragma unroll
        for (int c = 0; c < K; ++c) row_total += adjr32[p * K + c];
        static_cast<T *>(buf.node_adj)[mid_at * K + p] = static_cast<T>(row_total);
      }
      return;
    }
  }
  // FP64 form (every FP64 workspace; FP32 blocks outside the fast range).
  if (active) {
    if constexpr (!kFloat) {
#pragma unroll
      for (int m = 0; m < K; ++m) {
        const T lm = left[t][p * K + m];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] += lm * right[t][m * K + c0 + c];
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
      gr[t][p * K + c0 + c] = (a.v[c] != T(0) && acc[c] > T(0))
                                  ? static_cast<double>(a.v[c]) / static_cast<double>(acc[c])
                                  : 0.0;
  }
  __syncthreads();
  if (active) {
    // adj_L[p][m] for m = c0 .. c0 + 3
    Quad<T> q;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = c0 + i;
      double h = 0.0;
#pragma unroll
      for (int c = 0; c < K; ++c) h += gr[t][p * K + c] * static_cast<double>(right[t][m * K + c]);
      q.v[i] = static_cast<T>(static_cast<double>(left[t][p * K + m]) * h);
    }
    StoreAdjointQuad<T>(SlotAdjointTarget<T>(buf, d, op.left, b), static_cast<std::uint64_t>(p) * K + c0, q);
    // adj_R[m][c] for m = p, c = c0 .. c0 + 3
    const int m = p;
    double r[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int pp = 0; pp < K; ++pp) {
      const double lp = static_cast<double>(lraw[t][pp * K + m]);
#pragma unroll
      for (int c = 0; c < 4; ++c) r[c] += lp * gr[t][pp * K + c0 + c];
    }
    const double em = static_cast<double>(e_row);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double value = static_cast<double>(right[t][m * K + c0 + c]) * em * r[c];
      q.v[c] = static_cast<T>(value);
      adjr[t][m * K + c0 + c] = value;
    }
    StoreAdjointQuad<T>(SlotAdjointTarget<T>(buf, d, op.right, b), static_cast<std::uint64_t>(m) * K + c0, q);
  }
  __syncthreads();
  if (active && c0 == 0) {
    double total = 0.0;
#pragma unroll
    for (int c = 0; c < K; ++c) total += adjr[t][p * K + c];
    static_cast<T *>(buf.node_adj)[mid_at * K + p] = static_cast<T>(total);
  }
}

// log Z = scale(root) + log sum exp(root), in FP64 (src/inference.cc:778-784).
template <typename T>
__global__ void Finish(DeviceOps ops, Dims d, LinearBuffers buf, double *log_z) {
  const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (tid >= static_cast<std::uint64_t>(d.roots) * d.B) return;
  const std::uint32_t r = static_cast<std::uint32_t>(tid / d.B);
  const std::uint32_t b = static_cast<std::uint32_t>(tid % d.B);
  const std::uint64_t at = static_cast<std::uint64_t>(ops.roots[r]) * d.B + b;
  const T *x = static_cast<const T *>(buf.nodes) + at * d.K;
  double sum = 0.0;
  for (std::uint32_t s = 0; s < d.K; ++s) sum += exp(static_cast<double>(x[s]));
  log_z[static_cast<std::uint64_t>(b) * d.roots + r] = sum > 0.0 ? buf.node_scale[at] + log(sum) : kNegInf;
}

// ------------------------------------------------------------------------
// Reverse (adjoint)
// ------------------------------------------------------------------------

// Root adjoint = root marginal (src/inference.cc:338-343), times grad.
template <typename T>
__global__ void SeedAdjoint(DeviceOps ops, Dims d, LinearBuffers buf, const double *grad) {
  const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (tid >= static_cast<std::uint64_t>(d.roots) * d.B) return;
  const std::uint32_t r = static_cast<std::uint32_t>(tid / d.B);
  const std::uint32_t b = static_cast<std::uint32_t>(tid % d.B);
  const std::uint64_t at = static_cast<std::uint64_t>(ops.roots[r]) * d.B + b;
  const T *x = static_cast<const T *>(buf.nodes) + at * d.K;
  T *adj = static_cast<T *>(buf.node_adj) + at * d.K;
  double sum = 0.0;
  for (std::uint32_t s = 0; s < d.K; ++s) sum += exp(static_cast<double>(x[s]));
  const double g = grad ? grad[static_cast<std::uint64_t>(b) * d.roots + r] : 1.0;
  const double factor = sum > 0.0 ? g / sum : 0.0;
  for (std::uint32_t s = 0; s < d.K; ++s) adj[s] = static_cast<T>(exp(static_cast<double>(x[s])) * factor);
}

// Rake adjoint (src/inference.cc:488-518): with t[p][c] = P[p][c] exp(v[c])
// and m[p] = sum_c t[p][c], adj_P[p][c] = adj_m[p] t[p][c] / m[p] and
// adj_v[c] = sum_p adj_P[p][c]; the ratio adj_m / m is taken in FP64.
template <typename T, int G>
__global__ void RakeReverse(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, LinearBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const std::uint32_t r = offset + static_cast<std::uint32_t>(lane.item / d.B);
  const DeviceRake op = ops.rakes[r];
  const std::uint64_t leaf_at = static_cast<std::uint64_t>(op.leaf) * d.B + b;
  const T *v = static_cast<const T *>(buf.nodes) + leaf_at * d.K;
  const T *P = Slot<T>(buf, d, op.slot, b);
  const bool categorical = r < buf.cat.rakes;
  T own = T(0);
  if (categorical) {
    double unused;
    own = Exp(CategoricalLeaf<T, G>(buf.cat, d, op.leaf, b, lane.s, &unused));
  }
  // Lanes s >= K run the loops on row K - 1 and store nothing.
  const std::uint32_t s = lane.s < d.K ? lane.s : d.K - 1;
  const T *row = P + static_cast<std::uint64_t>(s) * d.K;
  T acc = T(0);
  for (std::uint32_t c = 0; c < d.K; ++c) acc += row[c] * LeafExp<T, G>(categorical, own, v, c);
  double g = 0.0;
  if (lane.s < d.K) {
    const T a = RakeAdjoint<T>(buf, d, op, b)[lane.s];
    g = (a != T(0) && acc > T(0)) ? static_cast<double>(a) / static_cast<double>(acc) : 0.0;
  }
  const AdjointTarget<T> out = SlotAdjointTarget<T>(buf, d, op.slot, b);
  const std::uint64_t at = static_cast<std::uint64_t>(s) * d.K;
  for (std::uint32_t c = 0; c < d.K; ++c) {
    const T e = LeafExp<T, G>(categorical, own, v, c);
    if (lane.s < d.K) StoreAdjointValue(out, at + c, static_cast<T>(g * static_cast<double>(row[c] * e)));
  }
  double column = 0.0;
  const T e_s = lane.s >= d.K ? T(0) : categorical ? own : Exp(v[lane.s]);
  for (std::uint32_t p = 0; p < d.K; ++p) {
    const double gp = __shfl(g, static_cast<int>(p), G);
    if (lane.s < d.K) column += gp * static_cast<double>(P[static_cast<std::uint64_t>(p) * d.K + lane.s] * e_s);
  }
  if (lane.s < d.K) static_cast<T *>(buf.node_adj)[leaf_at * d.K + lane.s] = static_cast<T>(column);
}

// exp of the K = 4 leaf vector of categorical node n, item b, as
// CategoricalLeaf forms it on the four lanes of a group and RakeReverse
// exponentiates it (the maximum taken over the four entries here; it is
// exact, so the order does not matter).
template <typename T>
__device__ inline void CategoricalLeafExp4(const CategoricalLeaves &c, const Dims &d, std::uint32_t n, std::uint32_t b,
                                           T e[4], double *scale) {
  const bool rooted = c.is_root[n] != 0;
  const std::uint32_t o = c.index[n];
  const std::uint32_t category = o == 0xFFFFFFFFu ? 0 : c.obs[static_cast<std::uint64_t>(o) * d.B + b];
  const T *root = static_cast<const T *>(c.root);
  double x[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    T value;
    if (o == 0xFFFFFFFFu) {
      value = rooted ? root[s] : T(1);
    } else {
      const T em = category < c.categories
                       ? static_cast<const T *>(c.emissions)[static_cast<std::uint64_t>(category) * 4 + s]
                       : T(0);
      value = (rooted ? root[s] : T(1)) * em;
    }
    x[s] = static_cast<double>(value);
  }
  if constexpr (std::is_same<T, float>::value) {
    if (c.element_logs) {
      float l[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) l[s] = ElementLog(static_cast<float>(x[s]));
      const float top = Max(Max(l[0], l[1]), Max(l[2], l[3]));
      const bool live = top > NegInf<float>();
      *scale = live ? static_cast<double>(top) : kNegInf;
#pragma unroll
      for (int s = 0; s < 4; ++s) e[s] = Exp(live ? l[s] - top : NegInf<float>());
      return;
    }
  }
  double l[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) l[s] = x[s] > 0.0 ? log(x[s]) : kNegInf;
  const double top = Max(Max(l[0], l[1]), Max(l[2], l[3]));
  *scale = top;
#pragma unroll
  for (int s = 0; s < 4; ++s) e[s] = Exp(top > kNegInf && x[s] > 0.0 ? static_cast<T>(l[s] - top) : NegInf<T>());
}

// The rake adjoint at K = 4 with one lane per (rake, item): RakeReverse's
// arithmetic in the same order (row sums in T, the ratio and the column sums
// in FP64), the slot rows, leaf vector and adjoint read as quads and each
// row of the slot adjoint stored as one quad, where the four-lane form
// repeats the leaf exponentials on every lane and stores single values.
template <typename T>
__global__ __launch_bounds__(256) void RakeReverseK4(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count,
                                                     LinearBuffers buf) {
  constexpr int K = 4;
  const std::uint64_t item = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (item >= static_cast<std::uint64_t>(count) * d.B) return;
  const std::uint32_t b = static_cast<std::uint32_t>(item % d.B);
  const std::uint32_t r = offset + static_cast<std::uint32_t>(item / d.B);
  const DeviceRake op = ops.rakes[r];
  const std::uint64_t leaf_at = static_cast<std::uint64_t>(op.leaf) * d.B + b;
  T e[K];
  if (r < buf.cat.rakes) {
    double unused;
    CategoricalLeafExp4<T>(buf.cat, d, op.leaf, b, e, &unused);
  } else {
    const Quad<T> v = LoadQuad(static_cast<const T *>(buf.nodes) + leaf_at * K);
#pragma unroll
    for (int c = 0; c < K; ++c) e[c] = Exp(v.v[c]);
  }
  const T *P = Slot<T>(buf, d, op.slot, b);
  Quad<T> row[K];
#pragma unroll
  for (int p = 0; p < K; ++p) row[p] = LoadQuad(P + p * K);
  const Quad<T> a = LoadQuad(RakeAdjoint<T>(buf, d, op, b));
  double g[K];
#pragma unroll
  for (int p = 0; p < K; ++p) {
    T acc = T(0);
#pragma unroll
    for (int c = 0; c < K; ++c) acc += row[p].v[c] * e[c];
    g[p] = (a.v[p] != T(0) && acc > T(0)) ? static_cast<double>(a.v[p]) / static_cast<double>(acc) : 0.0;
  }
  const AdjointTarget<T> out = SlotAdjointTarget<T>(buf, d, op.slot, b);
#pragma unroll
  for (int p = 0; p < K; ++p) {
    Quad<T> q;
#pragma unroll
    for (int c = 0; c < K; ++c) q.v[c] = static_cast<T>(g[p] * static_cast<double>(row[p].v[c] * e[c]));
    StoreAdjointQuad<T, false>(out, static_cast<std::uint64_t>(p) * K, q);
  }
  Quad<T> column;
#pragma unroll
  for (int c = 0; c < K; ++c) {
    double sum = 0.0;
#pragma unroll
    for (int p = 0; p < K; ++p) sum += g[p] * static_cast<double>(row[p].v[c] * e[c]);
    column.v[c] = static_cast<T>(sum);
  }
  StoreQuad(static_cast<T *>(buf.node_adj) + leaf_at * K, column);
}

// The rake forward at K = 4 with one lane per (rake, item): RakeForward's
// arithmetic in the same order, the slot rows and the leaf vector read as
// quads and the branch stored as one quad.
template <typename T>
__global__ __launch_bounds__(256) void RakeForwardK4(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count,
                                                     LinearBuffers buf) {
  constexpr int K = 4;
  const std::uint64_t item = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (item >= static_cast<std::uint64_t>(count) * d.B) return;
  const std::uint32_t b = static_cast<std::uint32_t>(item % d.B);
  const std::uint32_t r = offset + static_cast<std::uint32_t>(item / d.B);
  const DeviceRake op = ops.rakes[r];
  const std::uint64_t leaf_at = static_cast<std::uint64_t>(op.leaf) * d.B + b;
  T e[K];
  double leaf_scale;
  if (r < buf.cat.rakes) {
    CategoricalLeafExp4<T>(buf.cat, d, op.leaf, b, e, &leaf_scale);
  } else {
    const Quad<T> v = LoadQuad(static_cast<const T *>(buf.nodes) + leaf_at * K);
#pragma unroll
    for (int c = 0; c < K; ++c) e[c] = Exp(v.v[c]);
    leaf_scale = buf.node_scale[leaf_at];
  }
  const T *P = Slot<T>(buf, d, op.slot, b);
  double message[K];
#pragma unroll
  for (int p = 0; p < K; ++p) {
    const Quad<T> row = LoadQuad(P + p * K);
    T acc = T(0);
#pragma unroll
    for (int c = 0; c < K; ++c) acc += row.v[c] * e[c];
    message[p] = acc > T(0) ? RowScale(buf, d, op.slot, b, p) + log(static_cast<double>(acc)) : kNegInf;
  }
  const double top = Max(Max(message[0], message[1]), Max(message[2], message[3]));
  const std::uint64_t at = static_cast<std::uint64_t>(op.branch) * d.B + b;
  const bool live = top > kNegInf;
  Quad<T> q;
#pragma unroll
  for (int p = 0; p < K; ++p) q.v[p] = live ? static_cast<T>(message[p] - top) : NegInf<T>();
  StoreQuad(static_cast<T *>(buf.branches) + at * K, q);
  buf.branch_scale[at] = live ? leaf_scale + top : kNegInf;
}

// adj(branch) = adj(parent) (src/inference.cc:466-475)
template <typename T, int G>
__global__ void AbsorbReverse(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, LinearBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B || lane.s >= d.K) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const DeviceAbsorb op = ops.absorbs[offset + lane.item / d.B];
  static_cast<T *>(buf.branch_adj)[(static_cast<std::uint64_t>(op.branch) * d.B + b) * d.K + lane.s] =
      static_cast<const T *>(buf.node_adj)[(static_cast<std::uint64_t>(op.parent) * d.B + b) * d.K + lane.s];
}

// adj(source) = adj(destination) (src/inference.cc:477-486)
template <typename T, int G>
__global__ void CombineReverse(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, LinearBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B || lane.s >= d.K) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const DeviceCombine op = ops.combines[offset + lane.item / d.B];
  T *adj = static_cast<T *>(buf.branch_adj);
  adj[(static_cast<std::uint64_t>(op.source) * d.B + b) * d.K + lane.s] =
      adj[(static_cast<std::uint64_t>(op.destination) * d.B + b) * d.K + lane.s];
}

// Compression adjoint (src/inference.cc:416-464). With e the middle weights,
// S = L diag(e) R (the forward product before row normalisation) and
// Gr = adj_out / S (FP64; 0 where adj_out or S is 0):
//   adj_L = L diag(e) .* (Gr R^T),  adj_R = diag(e) R .* (L^T Gr),
//   adj_mid[m] = sum_c adj_R[m][c].
// Lane p forms row p of S, Gr and adj_L; Gr goes through LDS so that lane m
// can form row m of L^T Gr.
template <typename T, int G>
__global__ void CompressReverse(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, LinearBuffers buf) {
  extern __shared__ double shared_gr[];
  double *gr = shared_gr + (threadIdx.x / G) * G * G;
  const Lane lane = LaneOf<G>();
  const bool in_range = lane.item < static_cast<std::uint64_t>(count) * d.B;
  const bool active = in_range && lane.s < d.K;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  DeviceCompress op{};
  const T *L = nullptr;
  const T *R = nullptr;
  std::uint64_t mid_at = 0;
  T e[G];
  if (in_range) {
    op = ops.compresses[offset + lane.item / d.B];
    L = Slot<T>(buf, d, op.left, b);
    R = Slot<T>(buf, d, op.right, b);
    mid_at = static_cast<std::uint64_t>(op.middle) * d.B + b;
  }
  if (active) {
    MiddleWeights<T, G>(buf, d, op, b, static_cast<const T *>(buf.nodes) + mid_at * d.K, e);
    const std::uint64_t p = lane.s;
    T acc[G];
#pragma unroll
    for (int c = 0; c < G; ++c) acc[c] = T(0);
    for (std::uint32_t m = 0; m < d.K; ++m) {
      const T lm = L[p * d.K + m] * e[m];
      const T *row = R + static_cast<std::uint64_t>(m) * d.K;
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c < static_cast<int>(d.K)) acc[c] += lm * row[c];
    }
    const T *adj_out =
        static_cast<const T *>(buf.path_adj) + (static_cast<std::uint64_t>(op.out - d.E) * d.B + b) * d.K * d.K + p * d.K;
#pragma unroll
    for (int c = 0; c < G; ++c) {
      if (c < static_cast<int>(d.K)) {
        const T a = adj_out[c];
        gr[p * G + c] = (a != T(0) && acc[c] > T(0)) ? static_cast<double>(a) / static_cast<double>(acc[c]) : 0.0;
      }
    }
    const AdjointTarget<T> adj_left = SlotAdjointTarget<T>(buf, d, op.left, b);
    for (std::uint32_t m = 0; m < d.K; ++m) {
      const T *row = R + static_cast<std::uint64_t>(m) * d.K;
      double h = 0.0;
      for (std::uint32_t c = 0; c < d.K; ++c) h += gr[p * G + c] * static_cast<double>(row[c]);
      StoreAdjointValue(adj_left, p * d.K + m, static_cast<T>(static_cast<double>(L[p * d.K + m] * e[m]) * h));
    }
  }
  __syncthreads();
  if (active) {
    const std::uint64_t m = lane.s;
    double r[G];
#pragma unroll
    for (int c = 0; c < G; ++c) r[c] = 0.0;
    for (std::uint32_t p = 0; p < d.K; ++p) {
      const double lp = static_cast<double>(L[static_cast<std::uint64_t>(p) * d.K + m]);
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c < static_cast<int>(d.K)) r[c] += lp * gr[p * G + c];
    }
    const AdjointTarget<T> adj_right = SlotAdjointTarget<T>(buf, d, op.right, b);
    const T *row = R + m * d.K;
    const double em = static_cast<double>(e[m]);
    double total = 0.0;
#pragma unroll
    for (int c = 0; c < G; ++c) {
      if (c < static_cast<int>(d.K)) {
        const double value = static_cast<double>(row[c]) * em * r[c];
        StoreAdjointValue(adj_right, m * d.K + c, static_cast<T>(value));
        total += value;
      }
    }
    static_cast<T *>(buf.node_adj)[mid_at * d.K + m] = static_cast<T>(total);
  }
}

// ------------------------------------------------------------------------
// Staging
// ------------------------------------------------------------------------

// in [X][Y][K] -> out [Y][X][K] through a 64 x 64-element LDS tile. The grid
// is one-dimensional: block i covers x-tile i % x_tiles and y-tile
// i / x_tiles. Requires K <= 64.
//   kMode 0: plain copy (adjoints);
//   kMode 1: linear potentials -> max-shifted logs + FP64 scale;
//   kMode 2: log potentials -> max-shifted logs + FP64 scale.
template <typename In, typename Out, int kMode>
__global__ void TransposeVectors(const In *in, Out *out, std::uint32_t X, std::uint32_t Y, std::uint32_t K,
                                 double *scale, std::uint32_t x_tiles) {
  __shared__ double tile[64][65];
  const std::uint32_t ty = 64 / K;  // y-vectors per tile row
  const std::uint32_t row = ty * K; // elements per tile row
  // Block-stride over the tiles (a dispatch's grid size is a 32-bit count
  // of work-items); every thread of a block runs the same iterations.
  const std::uint64_t tiles = static_cast<std::uint64_t>(x_tiles) * ((Y + ty - 1) / ty);
  for (std::uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const std::uint64_t x0 = (t % x_tiles) * 64;
    const std::uint64_t y0 = (t / x_tiles) * ty;
    for (std::uint32_t i = threadIdx.x; i < 64 * row; i += blockDim.x) {
      const std::uint32_t x = i / row, j = i % row;
      const std::uint64_t gx = x0 + x;
      if (gx < X && y0 * K + j < static_cast<std::uint64_t>(Y) * K) {
        const double value = static_cast<double>(in[(gx * Y + y0) * K + j]);
        tile[x][j] = kMode == 1 ? (value > 0.0 ? log(value) : kNegInf) : value;
      }
    }
    __syncthreads();
    if (kMode == 0) {
      const std::uint32_t xk = 64 * K;
      for (std::uint32_t i = threadIdx.x; i < ty * xk; i += blockDim.x) {
        const std::uint32_t y = i / xk, r = i % xk;
        const std::uint32_t x = r / K, k = r % K;
        const std::uint64_t gy = y0 + y, gx = x0 + x;
        if (gy < Y && gx < X) out[(gy * X + gx) * K + k] = static_cast<Out>(tile[x][y * K + k]);
      }
    } else {
      // one thread per output vector
      for (std::uint32_t i = threadIdx.x; i < 64 * ty; i += blockDim.x) {
        const std::uint32_t y = i / 64, x = i % 64;
        const std::uint64_t gy = y0 + y, gx = x0 + x;
        if (gy >= Y || gx >= X) continue;
        double top = kNegInf;
        for (std::uint32_t k = 0; k < K; ++k) top = fmax(top, tile[x][y * K + k]);
        const bool live = top > kNegInf;
        Out *o = out + (gy * X + gx) * K;
        for (std::uint32_t k = 0; k < K; ++k) o[k] = live ? static_cast<Out>(tile[x][y * K + k] - top) : NegInf<Out>();
        scale[gy * X + gx] = top;
      }
    }
    __syncthreads();  // the tile is refilled by the next iteration
  }
}

// Log edge potentials [E][K][K] -> row-max-shifted exponentials + FP64 row
// scales [E][K].
template <typename T>
__global__ void InitEdgesFromLogs(const T *in, T *edges, double *scale, Dims d) {
  const std::uint64_t rows = static_cast<std::uint64_t>(d.E) * d.K;
  for (std::uint64_t tid = GridThread(); tid < rows; tid += GridStride()) {
    const T *x = in + tid * d.K;
    T top = NegInf<T>();
    for (std::uint32_t c = 0; c < d.K; ++c) top = Max(top, x[c]);
    const bool live = top > NegInf<T>();
    for (std::uint32_t c = 0; c < d.K; ++c)
      edges[tid * d.K + c] = live ? static_cast<T>(exp(static_cast<double>(x[c] - top))) : T(0);
    scale[tid] = live ? static_cast<double>(top) : kNegInf;
  }
}

// Categorical node factors (src/gpu_backend_impl.inc:326-382 semantics):
// (root ? root[s] : 1) * (observed ? (cat < C ? emission[cat][s] : 0) : 1),
// the product formed in the Scalar type P so that it is bitwise equal to the
// dense construction; stored raw (kLog false) or as max-shifted logs.
template <typename P, typename Out, bool kLog>
__global__ void InitCategorical(const std::uint8_t *obs, const std::uint32_t *obs_index, std::uint32_t obs_count,
                                std::uint32_t categories, const std::uint8_t *is_root, const P *root,
                                const P *emissions, Out *nodes, double *scale, Dims d,
                                const std::uint8_t *skip, bool element_logs) {
  const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (tid >= static_cast<std::uint64_t>(d.N) * d.B) return;
  const std::uint32_t n = static_cast<std::uint32_t>(tid / d.B);
  const std::uint32_t b = static_cast<std::uint32_t>(tid % d.B);
  if (skip != nullptr && skip[n] != 0) return;
  Out *out = nodes + tid * d.K;
  const bool rooted = is_root[n] != 0;
  const std::uint32_t o = obs_index[n];
  // obs is [observation node][batch] (transposed at upload): consecutive
  // threads read consecutive bytes.
  const std::uint32_t category = o == 0xFFFFFFFFu ? 0 : obs[static_cast<std::uint64_t>(o) * d.B + b];
  auto value = [&](std::uint32_t s) -> P {
    if (o == 0xFFFFFFFFu) return rooted ? root[s] : P(1);
    const P e = category < categories ? emissions[static_cast<std::uint64_t>(category) * d.K + s] : P(0);
    return (rooted ? root[s] : P(1)) * e;
  };
  if (!kLog) {
    for (std::uint32_t s = 0; s < d.K; ++s) out[s] = static_cast<Out>(value(s));
    return;
  }
  if (o == 0xFFFFFFFFu && !rooted) {  // factor 1: every log is 0, as is their maximum
    for (std::uint32_t s = 0; s < d.K; ++s) out[s] = Out(0);
    scale[tid] = 0.0;
    return;
  }
  if constexpr (std::is_same<P, float>::value && std::is_same<Out, float>::value) {
    if (element_logs) {  // as the team transpose: FP32 logs, shift and maximum
      float top = NegInf<float>();
      for (std::uint32_t s = 0; s < d.K; ++s) top = Max(top, ElementLog(value(s)));
      const bool live = top > NegInf<float>();
      for (std::uint32_t s = 0; s < d.K; ++s) out[s] = live ? ElementLog(value(s)) - top : NegInf<float>();
      scale[tid] = live ? static_cast<double>(top) : kNegInf;
      return;
    }
  }
  double top = kNegInf;
  for (std::uint32_t s = 0; s < d.K; ++s) {
    const double x = static_cast<double>(value(s));
    top = fmax(top, x > 0.0 ? log(x) : kNegInf);
  }
  const bool live = top > kNegInf;
  for (std::uint32_t s = 0; s < d.K; ++s) {
    const double x = static_cast<double>(value(s));
    out[s] = live && x > 0.0 ? static_cast<Out>(log(x) - top) : NegInf<Out>();
  }
  scale[tid] = top;
}

// sums[r] = sum over b of log_z[b * roots + r] in a fixed order: thread t
// of root r's block adds b = t, t + 256, ... in turn, then the 256 partial
// sums are added pairwise (t += t + 128, then t += t + 64, ..., 1).
constexpr unsigned kSumThreads = 256;
__global__ __launch_bounds__(kSumThreads) void SumOverBatch(const double *log_z, double *sums, Dims d) {
  __shared__ double part[kSumThreads];
  const std::uint32_t r = blockIdx.x, t = threadIdx.x;
  double sum = 0.0;
  for (std::uint32_t b = t; b < d.B; b += kSumThreads) sum += log_z[static_cast<std::uint64_t>(b) * d.roots + r];
  part[t] = sum;
  __syncthreads();
  for (unsigned s = kSumThreads / 2; s > 0; s >>= 1) {
    if (t < s) part[t] += part[t + s];
    __syncthreads();
  }
  if (t == 0) sums[r] = part[0];
}

template <typename T>
__global__ void ReduceEdges(const T *edge_out, T *out, Dims d) {
  const std::uint64_t kk = static_cast<std::uint64_t>(d.K) * d.K;
  const std::uint64_t total = static_cast<std::uint64_t>(d.E) * kk;
  for (std::uint64_t tid = GridThread(); tid < total; tid += GridStride()) {
    double sum = 0.0;
    for (std::uint32_t b = 0; b < d.B; ++b) sum += static_cast<double>(edge_out[b * total + tid]);
    out[tid] = static_cast<T>(sum);
  }
}

template <typename In, typename Out>
__global__ void Convert(const In *in, Out *out, std::uint64_t count) {
  for (std::uint64_t tid = GridThread(); tid < count; tid += GridStride()) out[tid] = static_cast<Out>(in[tid]);
}

// ------------------------------------------------------------------------
// Max-product (FP64 log domain)
// ------------------------------------------------------------------------

// Log factors (src/inference.cc:95-98: the log of a positive potential,
// else -inf): FP64 potentials through the device library's log, FP32 ones
// through LogOfFloat (log_f32.h) with its reduction table staged in LDS.
// Every MAP kernel logs an input through these, so a value's log has the
// same bits whichever kernel takes it.
__device__ const LogF32Entry kLogTableDevice[128] = THMM_LOG_F32_TABLE;

// Copies the reduction table into the block's LDS; every thread of the
// block calls it (before any early return).
__device__ inline void StageLogTable(LogF32Entry *table) {
  for (unsigned i = threadIdx.x; i < 128; i += blockDim.x) table[i] = kLogTableDevice[i];
  __syncthreads();
}

__device__ inline double LogFactor(double x, const LogF32Entry *) { return x > 0.0 ? log(x) : kNegInf; }
__device__ inline double LogFactor(float x, const LogF32Entry *table) {
  return x > 0.0f ? LogOfFloat(x, table) : kNegInf;
}

// Logs of public [B][N][K] node potentials into internal [N][B][K], and of
// the edges (src/inference.cc:95-98). With `ids`, only nodes ids[0 ..
// listed) are written.
template <typename In>
__global__ void MaxInit(const In *nodes_in, const In *edges_in, double *nodes, double *edges, Dims d,
                        const std::uint32_t *ids, std::uint32_t listed) {
  __shared__ LogF32Entry table[128];
  StageLogTable(table);
  const std::uint64_t node_values = static_cast<std::uint64_t>(ids != nullptr ? listed : d.N) * d.B * d.K;
  const std::uint64_t edge_values = static_cast<std::uint64_t>(d.E) * d.K * d.K;
  for (std::uint64_t tid = GridThread(); tid < node_values + edge_values; tid += GridStride()) {
    if (tid < node_values) {
      const std::uint64_t s = tid % d.K, nb = tid / d.K;
      const std::uint64_t b = nb % d.B, i = nb / d.B;
      const std::uint64_t n = ids != nullptr ? ids[i] : i;
      nodes[(n * d.B + b) * d.K + s] = LogFactor(nodes_in[(b * d.N + n) * d.K + s], table);
    } else {
      const std::uint64_t i = tid - node_values;
      edges[i] = LogFactor(edges_in[i], table);
    }
  }
}

// MaxInit four values at a time (K a multiple of 4, 16-byte aligned arrays):
// quad loads of the inputs, 32-byte stores of their FP64 logs, the same
// x > 0 ? log x : -inf per value.
template <typename In>
__global__ void MaxInitQuad(const In *nodes_in, const In *edges_in, double *nodes, double *edges, Dims d,
                            const std::uint32_t *ids, std::uint32_t listed) {
  __shared__ LogF32Entry table[128];
  StageLogTable(table);
  const std::uint64_t node_quads = static_cast<std::uint64_t>(ids != nullptr ? listed : d.N) * d.B * d.K / 4;
  const std::uint64_t edge_quads = static_cast<std::uint64_t>(d.E) * d.K * d.K / 4;
  const std::uint32_t kq = d.K / 4;
  for (std::uint64_t tid = GridThread(); tid < node_quads + edge_quads; tid += GridStride()) {
    Quad<In> x;
    double *out;
    if (tid < node_quads) {
      const std::uint64_t q = tid % kq, nb = tid / kq;
      const std::uint64_t b = nb % d.B, i = nb / d.B;
      const std::uint64_t n = ids != nullptr ? ids[i] : i;
      x = LoadQuad(nodes_in + (b * d.N + n) * d.K + 4 * q);
      out = nodes + (n * d.B + b) * d.K + 4 * q;
    } else {
      const std::uint64_t i = tid - node_quads;
      x = LoadQuad(edges_in + 4 * i);
      out = edges + 4 * i;
    }
    Quad<double> y;
#pragma unroll
    for (int r = 0; r < 4; ++r) y.v[r] = LogFactor(x.v[r], table);
    StoreQuad(out, y);
  }
}

__device__ inline const double *MaxSlot(const MaxBuffers &buf, const Dims &d, std::uint32_t slot, std::uint32_t b) {
  const std::uint64_t kk = static_cast<std::uint64_t>(d.K) * d.K;
  return slot < d.E ? buf.edges + slot * kk : buf.paths + (static_cast<std::uint64_t>(slot - d.E) * d.B + b) * kk;
}

// src/inference.cc:576-594: strict '>' argmax of path + leaf.
template <int G>
__global__ void MaxRake(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, MaxBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B || lane.s >= d.K) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const DeviceRake op = ops.rakes[offset + lane.item / d.B];
  const double *v = buf.nodes + (static_cast<std::uint64_t>(op.leaf) * d.B + b) * d.K;
  const double *row = MaxSlot(buf, d, op.slot, b) + static_cast<std::uint64_t>(lane.s) * d.K;
  double best = row[0] + v[0];
  std::uint32_t arg = 0;
  for (std::uint32_t c = 1; c < d.K; ++c) {
    const double candidate = row[c] + v[c];
    if (candidate > best) {
      best = candidate;
      arg = c;
    }
  }
  const std::uint64_t at = (static_cast<std::uint64_t>(op.branch) * d.B + b) * d.K + lane.s;
  buf.branches[at] = best;
  buf.rake_choice[at] = static_cast<std::uint8_t>(arg);
}

template <int G>
__global__ void MaxCombine(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, MaxBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B || lane.s >= d.K) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const DeviceCombine op = ops.combines[offset + lane.item / d.B];
  buf.branches[(static_cast<std::uint64_t>(op.destination) * d.B + b) * d.K + lane.s] +=
      buf.branches[(static_cast<std::uint64_t>(op.source) * d.B + b) * d.K + lane.s];
}

// buf.raw_leaves: round 0's absorption, the parent's first write, whose log
// potential is taken here (the reduction table read from global memory).
template <int G>
__global__ void MaxAbsorb(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, MaxBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B || lane.s >= d.K) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const DeviceAbsorb op = ops.absorbs[offset + lane.item / d.B];
  double *x = buf.nodes + (static_cast<std::uint64_t>(op.parent) * d.B + b) * d.K + lane.s;
  double base;
  if (buf.raw_leaves != nullptr) {
    const std::uint64_t at = (static_cast<std::uint64_t>(b) * d.N + op.parent) * d.K + lane.s;
    base = buf.raw_leaves_f32 ? LogFactor(static_cast<const float *>(buf.raw_leaves)[at], kLogTableDevice)
                              : LogFactor(static_cast<const double *>(buf.raw_leaves)[at], nullptr);
  } else {
    base = *x;
  }
  *x = base + buf.branches[(static_cast<std::uint64_t>(op.branch) * d.B + b) * d.K + lane.s];
}

// src/inference.cc:614-639: candidate (L + mid) + R, strict '>'.
template <int G>
__global__ void MaxCompress(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, MaxBuffers buf) {
  const Lane lane = LaneOf<G>();
  if (lane.item >= static_cast<std::uint64_t>(count) * d.B || lane.s >= d.K) return;
  const std::uint32_t b = static_cast<std::uint32_t>(lane.item % d.B);
  const DeviceCompress op = ops.compresses[offset + lane.item / d.B];
  const double *L = MaxSlot(buf, d, op.left, b) + static_cast<std::uint64_t>(lane.s) * d.K;
  const double *R = MaxSlot(buf, d, op.right, b);
  const double *mid = buf.nodes + (static_cast<std::uint64_t>(op.middle) * d.B + b) * d.K;
  const std::uint64_t out_at = static_cast<std::uint64_t>(op.out - d.E) * d.B + b;
  double *out = buf.paths + out_at * d.K * d.K + static_cast<std::uint64_t>(lane.s) * d.K;
  std::uint8_t *choice = buf.compress_choice + out_at * d.K * d.K + static_cast<std::uint64_t>(lane.s) * d.K;
  for (std::uint32_t c = 0; c < d.K; ++c) {
    double best = (L[0] + mid[0]) + R[c];
    std::uint32_t arg = 0;
    for (std::uint32_t m = 1; m < d.K; ++m) {
      const double candidate = (L[m] + mid[m]) + R[static_cast<std::uint64_t>(m) * d.K + c];
      if (candidate > best) {
        best = candidate;
        arg = m;
      }
    }
    out[c] = best;
    choice[c] = static_cast<std::uint8_t>(arg);
  }
}

// Team forms of MaxRake and MaxCompress for K = 8 and 16 (FP64 state, the
// same association and strict '>' argmax, so bit-identical results): a
// lane owns a quad of four consecutive columns of one row, loads and stores
// move whole quads, and choices are stored four at a time.


// Four consecutive log factors of a slot from element `at` (a multiple of
// 4): edge slots from the raw potentials when given, logged here.
__device__ inline Quad<double> MaxQuad(const MaxBuffers &buf, const Dims &d, std::uint32_t slot, std::uint32_t b,
                                       std::uint64_t at, const LogF32Entry *table) {
  if (slot < d.E && buf.raw_edges != nullptr) {
    const std::uint64_t i = static_cast<std::uint64_t>(slot) * d.K * d.K + at;
    Quad<double> q;
    if (buf.raw_edges_f32) {
      const Quad<float> f = LoadQuad(static_cast<const float *>(buf.raw_edges) + i);
#pragma unroll
      for (int r = 0; r < 4; ++r) q.v[r] = LogFactor(f.v[r], table);
    } else {
      const Quad<double> f = LoadQuad(static_cast<const double *>(buf.raw_edges) + i);
#pragma unroll
      for (int r = 0; r < 4; ++r) q.v[r] = LogFactor(f.v[r], table);
    }
    return q;
  }
  return LoadQuad(MaxSlot(buf, d, slot, b) + at);
}

// The leaf's log quad at columns c0 .. c0 + 3 for a rake team of kT lanes
// whose rows take kQ lanes each: from the node state, or (buf.raw_leaves,
// the first rake level) logged here from the caller's potentials as MaxInit
// logs them. The team's rows repeat the same quads, so lane l logs value
// (l / kQ) % 4 of its quad and takes the other three from the lanes that
// logged them.
template <int K, int kQ, int kT>
__device__ inline Quad<double> MaxLeafQuad(const MaxBuffers &buf, const Dims &d, std::uint32_t leaf, std::uint32_t b,
                                           int c0, int l, const LogF32Entry *table) {
  static_assert(kT / kQ >= 4, "four rows per team");
  if (buf.raw_leaves == nullptr) return LoadQuad(buf.nodes + (static_cast<std::uint64_t>(leaf) * d.B + b) * K + c0);
  const std::uint64_t at = (static_cast<std::uint64_t>(b) * d.N + leaf) * K + c0 + (l / kQ) % 4;
  const double mine = buf.raw_leaves_f32 ? LogFactor(static_cast<const float *>(buf.raw_leaves)[at], table)
                                         : LogFactor(static_cast<const double *>(buf.raw_leaves)[at], table);
  Quad<double> v;
#pragma unroll
  for (int i = 0; i < 4; ++i) v.v[i] = __shfl(mine, l % kQ + kQ * i, kT);
  return v;
}

// The absorption of a rake whose branch is absorbed as it stands, in the
// rake kernel: entry p of the parent's node vector plus the rake's value for
// it (MaxAbsorb's arithmetic). With buf.raw_leaves (round 0) the parent's
// log potential is taken here, as MaxAbsorb takes it there.
__device__ inline void AbsorbDirect(const MaxBuffers &buf, const Dims &d, std::uint32_t parent, std::uint32_t b,
                                    std::uint32_t K, std::uint32_t p, double value, const LogF32Entry *table) {
  double *x = buf.nodes + (static_cast<std::uint64_t>(parent) * d.B + b) * K + p;
  double base;
  if (buf.raw_leaves != nullptr) {
    const std::uint64_t at = (static_cast<std::uint64_t>(b) * d.N + parent) * K + p;
    base = buf.raw_leaves_f32 ? LogFactor(static_cast<const float *>(buf.raw_leaves)[at], table)
                              : LogFactor(static_cast<const double *>(buf.raw_leaves)[at], table);
  } else {
    base = *x;
  }
  *x = base + value;
}

// Combines two (value, index) argmax partials of one row: the larger value,
// and on a tie the lower index, as an ascending strict '>' scan does.
__device__ inline void ArgMaxMerge(double &best, std::uint32_t &arg, double other, std::uint32_t other_arg) {
  if (other > best || (other == best && other_arg < arg)) best = other, arg = other_arg;
}

// Linear-domain screen of one rake row of raw edge factors f against the
// leaf's log vector v (kQ consecutive lanes hold the row's K columns, 4
// each; ScreenLeaf then ScreenRow): the mask of this lane's columns c whose
// log-domain value log f_c + v_c may be the row's largest, so that only
// those are logged.
//
// l_c = f_c 2^y_c, y_c = (v_c - vmax) log2 e, is formed in FP32: y_c from
// the FP64 difference rounded to FP32 (absolute error below 2^-15 where
// y_c >= -254, the only columns with l_c >= 2^-126 when f <= 2^126), then
// ldexp(f_c exp2(y_c - n), n) with n = floor(y_c) and the hardware exp2.
// Normal values carry a relative error below 2^-15, and no weight
// underflows on its own: an l_c flushed to 0 is below 2^-126 in exact
// arithmetic. With lmax the row's largest l_c >= 2^-100, a column below
// lmax (1 - 2^-12) has log f_c + v_c below the winner's by more than 2^-13
// in exact arithmetic. The FP64 rounding of either sum (below 2^-21 at
// |v| < 2^30) cannot close that gap: it neither wins nor ties. Rows outside
// these conditions log every column: vmax not finite or |vmax| >= 2^30,
// a factor above 2^126 or not a number, or lmax < 2^-100.

// The leaf-side half of the screen, shared by every row of a rake: vmax
// over the row's kQ lanes (every lane of a row holds four of the K
// columns), and for each of the lane's columns n_c = floor(y_c) and
// 2^(y_c - n_c) where y_c >= -300 (ok_c), else nothing.
struct ScreenColumns {
  bool bounded;  // vmax finite and |vmax| < 2^30
  bool ok[4];
  float scale[4];
  int shift[4];
};

template <int kQ>
__device__ inline ScreenColumns ScreenLeaf(const Quad<double> &v) {
  double vmax = fmax(fmax(v.v[0], v.v[1]), fmax(v.v[2], v.v[3]));
#pragma unroll
  for (int o = 1; o < kQ; o <<= 1) vmax = fmax(vmax, __shfl_xor(vmax, o, kQ));
  ScreenColumns s;
  s.bounded = vmax > kNegInf && fabs(vmax) < 0x1p30;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float y = static_cast<float>(v.v[c] - vmax) * 1.44269504f;
    s.ok[c] = s.bounded && y > -300.0f;  // false for -inf and NaN
    const float n = s.ok[c] ? floorf(y) : 0.0f;
    s.scale[c] = s.ok[c] ? __builtin_amdgcn_exp2f(y - n) : 0.0f;
    s.shift[c] = static_cast<int>(n);
  }
  return s;
}

// The row half: l_c = ldexp(f_c 2^(y_c - n_c), n_c) for the lane's columns
// and the mask of those that may win (above).
template <int kQ, typename F>
__device__ inline unsigned ScreenRow(const ScreenColumns &s, const F (&f)[4]) {
  // 0 <= f <= 2^126 for every column, NaN included (a failed comparison).
  float wild = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; ++c) wild = (f[c] >= F(0) && f[c] <= F(0x1p126)) ? wild : 1.0f;
#pragma unroll
  for (int o = 1; o < kQ; o <<= 1) wild = fmaxf(wild, __shfl_xor(wild, o, kQ));
  const bool bounded = s.bounded && wild == 0.0f;
  float lin[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
    lin[c] = bounded && s.ok[c] ? ldexpf(static_cast<float>(f[c]) * s.scale[c], s.shift[c]) : 0.0f;
  float lmax = fmaxf(fmaxf(lin[0], lin[1]), fmaxf(lin[2], lin[3]));
#pragma unroll
  for (int o = 1; o < kQ; o <<= 1) lmax = fmaxf(lmax, __shfl_xor(lmax, o, kQ));
  if (!bounded || !(lmax >= 0x1p-100f)) return 0xFu;
  const float bar = lmax * (1.0f - 0x1p-12f);
  unsigned mask = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c)
    if (lin[c] >= bar) mask |= 1u << c;
  return mask;
}

// One rake row per kQ = K/4 consecutive lanes; K rows per (op, item).
// Raw edge factors are screened (ScreenLeaf, ScreenRow) before they are logged;
// the values compared are the same FP64 sums as without the screen.
template <int K>
__global__ __launch_bounds__(256) void MaxRakeTeam(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count,
                                                   MaxBuffers buf) {
  constexpr int kQ = K / 4, kLanes = K * kQ;
  __shared__ LogF32Entry table[128];
  StageLogTable(table);
  const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const std::uint64_t item = tid / kLanes;
  const int l = static_cast<int>(tid % kLanes);
  if (item >= static_cast<std::uint64_t>(count) * d.B) return;  // whole teams
  const std::uint32_t b = static_cast<std::uint32_t>(item % d.B);
  const DeviceRake op = ops.rakes[offset + item / d.B];
  const int p = l / kQ, c0 = (l % kQ) * 4;
  const Quad<double> v = MaxLeafQuad<K, kQ, kLanes>(buf, d, op.leaf, b, c0, l, table);
  double best = kNegInf;  // the strict '>' scan from -inf: a row of -inf keeps its first column
  std::uint32_t arg = c0;
  // The sequential scan starts at column 0's candidate (src/inference.cc:
  // 113-126): when that candidate is NaN, no later one compares greater and
  // the row keeps (NaN, 0). The merge below starts from -inf, so that case is
  // taken over explicitly after it.
  bool nan_first = false;
  if (op.slot < d.E && buf.raw_edges != nullptr) {
    const std::uint64_t at = static_cast<std::uint64_t>(op.slot) * K * K + p * K + c0;
    double f[4];
    if (buf.raw_edges_f32) {
      const Quad<float> q = LoadQuad(static_cast<const float *>(buf.raw_edges) + at);
#pragma unroll
      for (int c = 0; c < 4; ++c) f[c] = static_cast<double>(q.v[c]);
    } else {
      const Quad<double> q = LoadQuad(static_cast<const double *>(buf.raw_edges) + at);
#pragma unroll
      for (int c = 0; c < 4; ++c) f[c] = q.v[c];
    }
    // log f + v is NaN when v is NaN, or when log f = -inf (f <= 0 or NaN)
    // meets v = +inf, or f = +inf meets v = -inf.
    if (c0 == 0)
      nan_first = isnan(v.v[0]) || (!(f[0] > 0.0) && v.v[0] == __builtin_huge_val()) ||
                  (f[0] == __builtin_huge_val() && v.v[0] == kNegInf);
    // Ascending columns: the strict '>' keeps the lower index on a tie.
    for (unsigned mask = ScreenRow<kQ>(ScreenLeaf<kQ>(v), f); mask != 0; mask &= mask - 1) {
      const int c = __builtin_ctz(mask);
      // An FP32 factor is logged as the float it is (an exact round trip).
      const double log_f =
          buf.raw_edges_f32 ? LogFactor(static_cast<float>(f[c]), table) : LogFactor(f[c], table);
      const double candidate = log_f + v.v[c];
      if (candidate > best) best = candidate, arg = c0 + c;
    }
  } else {
    const Quad<double> row = LoadQuad(MaxSlot(buf, d, op.slot, b) + p * K + c0);
    if (c0 == 0) nan_first = isnan(row.v[0] + v.v[0]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double candidate = row.v[c] + v.v[c];
      if (candidate > best) best = candidate, arg = c0 + c;
    }
  }
#pragma unroll
  for (int o = 1; o < kQ; o <<= 1) {
    const double other = __shfl_xor(best, o, kQ);
    const std::uint32_t other_arg = __shfl_xor(arg, o, kQ);
    ArgMaxMerge(best, arg, other, other_arg);
  }
  if (nan_first) best = __builtin_nan(""), arg = 0;
  // The choices of rows p .. p + 3 (p % 4 == 0) go out as one 4-byte store
  // (lanes c0 == 0 of the next three rows are kQ, 2 kQ, 3 kQ lanes on).
  std::uint32_t packed = arg;
#pragma unroll
  for (int i = 1; i < 4; ++i)
    packed |= static_cast<std::uint32_t>(__shfl(arg, (l + i * kQ) % kLanes, kLanes)) << (8 * i);
  if (c0 == 0) {
    const std::uint64_t at = (static_cast<std::uint64_t>(op.branch) * d.B + b) * K + p;
    if (buf.rake_direct != nullptr && buf.rake_direct[offset + item / d.B] != 0)
      AbsorbDirect(buf, d, op.parent, b, K, p, best, table);
    else
      buf.branches[at] = best;
    if (p % 4 == 0) *reinterpret_cast<std::uint32_t *>(buf.rake_choice + at) = packed;
  }
}

// MaxRakeTeam over row steps, K = 16 and 32: kT lanes per (op, item)
// (16 at K = 16, 64 at K = 32) take kRps = 4 kT / K rows per step over
// kJ = K / kRps steps, each row on kQ = K / 4 consecutive lanes as above.
// A lane keeps the same four columns over its rows, so the leaf quad is
// loaded once and the screen's leaf half is formed once per lane; each row
// is screened, logged and merged exactly as in MaxRakeTeam (the same values
// and choices). After the merge every lane of a row holds its result; the
// first K lanes store the rows in order.
template <int K>
__global__ __launch_bounds__(256) void MaxRakeSteps(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count,
                                                    MaxBuffers buf) {
  constexpr int kQ = K / 4, kT = K <= 16 ? 16 : 64, kRps = 4 * kT / K, kJ = K / kRps;
  static_assert(kJ <= kQ && kRps * kQ == kT, "row steps");
  __shared__ LogF32Entry table[128];
  StageLogTable(table);
  const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const std::uint64_t item = tid / kT;
  const int l = static_cast<int>(tid % kT);
  if (item >= static_cast<std::uint64_t>(count) * d.B) return;  // whole teams
  const std::uint32_t b = static_cast<std::uint32_t>(item % d.B);
  const DeviceRake op = ops.rakes[offset + item / d.B];
  const int p0 = l / kQ, c0 = (l % kQ) * 4;
  const Quad<double> v = MaxLeafQuad<K, kQ, kT>(buf, d, op.leaf, b, c0, l, table);
  double best[kJ];
  std::uint32_t arg[kJ];
  bool nan_first[kJ];
#pragma unroll
  for (int j = 0; j < kJ; ++j) best[j] = kNegInf, arg[j] = c0, nan_first[j] = false;
  if (op.slot < d.E && buf.raw_edges != nullptr) {
    const std::uint64_t at = static_cast<std::uint64_t>(op.slot) * K * K + p0 * K + c0;
    const ScreenColumns screen = ScreenLeaf<kQ>(v);
    if (buf.raw_edges_f32) {
      Quad<float> q[kJ];
#pragma unroll
      for (int j = 0; j < kJ; ++j) q[j] = LoadQuad(static_cast<const float *>(buf.raw_edges) + at + j * kRps * K);
#pragma unroll
      for (int j = 0; j < kJ; ++j) {
        // A four-value copy per row: indexed by column below, it stays in
        // registers (an index into q would place q in scratch memory).
        float f[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) f[c] = q[j].v[c];
        if (c0 == 0)
          nan_first[j] = isnan(v.v[0]) || (!(f[0] > 0.0f) && v.v[0] == __builtin_huge_val()) ||
                         (f[0] == __builtin_huge_valf() && v.v[0] == kNegInf);
        for (unsigned mask = ScreenRow<kQ>(screen, f); mask != 0; mask &= mask - 1) {
          const int c = __builtin_ctz(mask);
          const double candidate = LogFactor(f[c], table) + v.v[c];
          if (candidate > best[j]) best[j] = candidate, arg[j] = c0 + c;
        }
      }
    } else {
      Quad<double> q[kJ];
#pragma unroll
      for (int j = 0; j < kJ; ++j) q[j] = LoadQuad(static_cast<const double *>(buf.raw_edges) + at + j * kRps * K);
#pragma unroll
      for (int j = 0; j < kJ; ++j) {
        double f[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) f[c] = q[j].v[c];
        if (c0 == 0)
          nan_first[j] = isnan(v.v[0]) || (!(f[0] > 0.0) && v.v[0] == __builtin_huge_val()) ||
                         (f[0] == __builtin_huge_val() && v.v[0] == kNegInf);
        for (unsigned mask = ScreenRow<kQ>(screen, f); mask != 0; mask &= mask - 1) {
          const int c = __builtin_ctz(mask);
          const double candidate = LogFactor(f[c], table) + v.v[c];
          if (candidate > best[j]) best[j] = candidate, arg[j] = c0 + c;
        }
      }
    }
  } else {
    const double *slot = MaxSlot(buf, d, op.slot, b) + p0 * K + c0;
    Quad<double> row[kJ];
#pragma unroll
    for (int j = 0; j < kJ; ++j) row[j] = LoadQuad(slot + j * kRps * K);
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
      if (c0 == 0) nan_first[j] = isnan(row[j].v[0] + v.v[0]);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double candidate = row[j].v[c] + v.v[c];
        if (candidate > best[j]) best[j] = candidate, arg[j] = c0 + c;
      }
    }
  }
  double out = 0.0;
  std::uint32_t out_arg = 0;
  const int q = l % kQ;
#pragma unroll
  for (int j = 0; j < kJ; ++j) {
    int nan_row = nan_first[j] ? 1 : 0;
#pragma unroll
    for (int o = 1; o < kQ; o <<= 1) {
      const double other = __shfl_xor(best[j], o, kQ);
      const std::uint32_t other_arg = __shfl_xor(arg[j], o, kQ);
      ArgMaxMerge(best[j], arg[j], other, other_arg);
      nan_row |= __shfl_xor(nan_row, o, kQ);
    }
    if (nan_row) best[j] = __builtin_nan(""), arg[j] = 0;
    if (j == q) out = best[j], out_arg = arg[j];
  }
  // Lane (p0, q) holds row q kRps + p0. Lane r < K takes row r from lane
  // (r % kRps) kQ + r / kRps so that the stores run in lane order: scattered
  // lanes would each cost a partial-line write to HBM. The choices of rows
  // r .. r + 3 (r % 4 == 0) go out as one 4-byte store.
  const int source = (l % kRps) * kQ + (l / kRps) % kJ;
  const double row_best = __shfl(out, source, kT);
  const std::uint32_t row_arg = __shfl(out_arg, source, kT);
  std::uint32_t packed = row_arg;
#pragma unroll
  for (int i = 1; i < 4; ++i) packed |= static_cast<std::uint32_t>(__shfl(row_arg, (l + i) % kT, kT)) << (8 * i);
  if (l < K) {
    const std::uint64_t at = (static_cast<std::uint64_t>(op.branch) * d.B + b) * K + l;
    if (buf.rake_direct != nullptr && buf.rake_direct[offset + item / d.B] != 0)
      AbsorbDirect(buf, d, op.parent, b, K, l, row_best, table);
    else
      buf.branches[at] = row_best;
    if (l % 4 == 0) *reinterpret_cast<std::uint32_t *>(buf.rake_choice + at) = packed;
  }
}

// K*K/4 lanes per (op, item), 256 / (K*K/4) items per block: L + mid and R
// are staged in LDS with quad loads, then each lane scans the K middle
// states for its quad of outputs.
template <int K>
__global__ __launch_bounds__(256) void MaxCompressTeam(DeviceOps ops, Dims d, std::uint32_t offset,
                                                       std::uint32_t count, MaxBuffers buf) {
  constexpr int kQ = K / 4, kLanes = K * kQ, kItems = 256 / kLanes;
  __shared__ double left_mid[kItems][K * K];  // L[p][m] + mid[m]
  __shared__ double right[kItems][K * K];
  __shared__ LogF32Entry table[128];
  StageLogTable(table);
  const int t = threadIdx.x / kLanes, l = threadIdx.x % kLanes;
  const std::uint64_t item = static_cast<std::uint64_t>(blockIdx.x) * kItems + t;
  const bool active = item < static_cast<std::uint64_t>(count) * d.B;
  DeviceCompress op{};
  std::uint32_t b = 0;
  if (active) {
    b = static_cast<std::uint32_t>(item % d.B);
    op = ops.compresses[offset + item / d.B];
    const double *mid = buf.nodes + (static_cast<std::uint64_t>(op.middle) * d.B + b) * K;
    const Quad<double> lq = MaxQuad(buf, d, op.left, b, 4 * l, table), rq = MaxQuad(buf, d, op.right, b, 4 * l, table);
    const Quad<double> mq = LoadQuad(mid + (4 * l) % K);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      left_mid[t][4 * l + i] = lq.v[i] + mq.v[i];
      right[t][4 * l + i] = rq.v[i];
    }
  }
  __syncthreads();
  if (!active) return;
  const int p = l / kQ, c0 = (l % kQ) * 4;
  double best[4];
  std::uint32_t arg[4] = {0, 0, 0, 0};
  const double lm0 = left_mid[t][p * K];
#pragma unroll
  for (int c = 0; c < 4; ++c) best[c] = lm0 + right[t][c0 + c];
#pragma unroll
  for (int m = 1; m < K; ++m) {
    const double lm = left_mid[t][p * K + m];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double candidate = lm + right[t][m * K + c0 + c];
      if (candidate > best[c]) best[c] = candidate, arg[c] = m;
    }
  }
  const std::uint64_t out_at = static_cast<std::uint64_t>(op.out - d.E) * d.B + b;
  Quad<double> q;
#pragma unroll
  for (int c = 0; c < 4; ++c) q.v[c] = best[c];
  StoreQuad(buf.paths + out_at * K * K + p * K + c0, q);
  const std::uint32_t packed = arg[0] | (arg[1] << 8) | (arg[2] << 16) | (arg[3] << 24);
  *reinterpret_cast<std::uint32_t *>(buf.compress_choice + out_at * K * K + p * K + c0) = packed;
}

// One lane's quad of a max-product compression from staged L + mid and R
// (the scan of MaxCompressTeam: row l / (K/4), columns 4 (l % (K/4)) .. +3,
// candidates (L + mid) + R over ascending m with strict '>'); the quad is
// elements 4l .. 4l+3 of the output path, the quad lane l loads of it as an
// operand. Stores the lane's four choices into `choices`, the compression's
// K x K choice array of this item.
template <int K>
__device__ inline Quad<double> MaxScanQuad(const double *left_mid, const double *right, int l,
                                           std::uint8_t *choices) {
  constexpr int kQ = K / 4;
  const int p = l / kQ, c0 = (l % kQ) * 4;
  Quad<double> best;
  std::uint32_t arg[4] = {0, 0, 0, 0};
  const double lm0 = left_mid[p * K];
#pragma unroll
  for (int c = 0; c < 4; ++c) best.v[c] = lm0 + right[c0 + c];
#pragma unroll
  for (int m = 1; m < K; ++m) {
    const double lm = left_mid[p * K + m];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double candidate = lm + right[m * K + c0 + c];
      if (candidate > best.v[c]) best.v[c] = candidate, arg[c] = m;
    }
  }
  *reinterpret_cast<std::uint32_t *>(choices + 4 * l) = arg[0] | (arg[1] << 8) | (arg[2] << 16) | (arg[3] << 24);
  return best;
}

// Max-product compression pieces (DeviceMaxPiece), K*K/4 lanes per (piece,
// item) as in MaxCompressTeam: the absorbed input compressions are evaluated
// first, one after the other, their paths kept in the lanes' registers (lane l's
// output quad is the quad it stages of an operand), then the root. Every
// compression is the scan of MaxCompressTeam on the same FP64 values, so
// paths and choices are bit-identical to the level-by-level form; only the
// absorbed paths are never written or read back. The absorbed compressions'
// middles are final when the root's level runs (a compression's middle node
// leaves the tree), and their inputs are slots written at earlier levels.
template <int K>
__global__ __launch_bounds__(256) void MaxPieceTeam(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count,
                                                    MaxBuffers buf) {
  constexpr int kQ = K / 4, kLanes = K * kQ, kItems = 256 / kLanes;
  // L + mid and R of one compression at a time: the absorbed inputs in
  // turn, then the root (half the LDS of staging both inputs at once, so
  // more wavefronts fit per SIMD).
  __shared__ double stage[kItems][2][K * K];
  __shared__ LogF32Entry table[128];
  StageLogTable(table);
  const int t = threadIdx.x / kLanes, l = threadIdx.x % kLanes;
  const std::uint64_t item = static_cast<std::uint64_t>(blockIdx.x) * kItems + t;
  const bool active = item < static_cast<std::uint64_t>(count) * d.B;
  DeviceCompress root{};
  std::uint32_t b = 0, flags = 0;
  Quad<double> side[2], root_mid;
  if (active) {
    b = static_cast<std::uint32_t>(item % d.B);
    const DeviceMaxPiece piece = ops.max_pieces[offset + item / d.B];
    flags = piece.flags;
    root = ops.compresses[piece.root];
  }
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const bool absorbed = active && (flags & (1u << s)) != 0;
    const std::uint32_t slot = s == 0 ? root.left : root.right;
    if (absorbed) {
      const DeviceCompress op = ops.compresses[slot - d.E];
      const Quad<double> lq = MaxQuad(buf, d, op.left, b, 4 * l, table),
                         rq = MaxQuad(buf, d, op.right, b, 4 * l, table);
      const Quad<double> mq =
          LoadQuad(buf.nodes + (static_cast<std::uint64_t>(op.middle) * d.B + b) * K + (4 * l) % K);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        stage[t][0][4 * l + i] = lq.v[i] + mq.v[i];
        stage[t][1][4 * l + i] = rq.v[i];
      }
    }
    __syncthreads();
    if (absorbed)
      side[s] = MaxScanQuad<K>(stage[t][0], stage[t][1], l,
                               buf.compress_choice + (static_cast<std::uint64_t>(slot - d.E) * d.B + b) * K * K);
    __syncthreads();
  }
  if (active) {
    // The root's slot operands and middle are loaded here, after the
    // absorbed inputs, which keeps the kernel at 72 registers.
    root_mid = LoadQuad(buf.nodes + (static_cast<std::uint64_t>(root.middle) * d.B + b) * K + (4 * l) % K);
#pragma unroll
    for (int s = 0; s < 2; ++s)
      if (!(flags & (1u << s))) side[s] = MaxQuad(buf, d, s == 0 ? root.left : root.right, b, 4 * l, table);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      stage[t][0][4 * l + i] = side[0].v[i] + root_mid.v[i];
      stage[t][1][4 * l + i] = side[1].v[i];
    }
  }
  __syncthreads();
  if (!active) return;
  const std::uint64_t out_at = static_cast<std::uint64_t>(root.out - d.E) * d.B + b;
  const Quad<double> q = MaxScanQuad<K>(stage[t][0], stage[t][1], l, buf.compress_choice + out_at * K * K);
  StoreQuad(buf.paths + out_at * K * K + 4 * l, q);
}

// src/inference.cc:641-649
__global__ void MaxFinish(DeviceOps ops, Dims d, MaxBuffers buf, double *log_weight, std::uint32_t *assignment) {
  const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (tid >= static_cast<std::uint64_t>(d.roots) * d.B) return;
  const std::uint32_t r = static_cast<std::uint32_t>(tid / d.B);
  const std::uint32_t b = static_cast<std::uint32_t>(tid % d.B);
  const std::uint32_t root = ops.roots[r];
  const double *x = buf.nodes + (static_cast<std::uint64_t>(root) * d.B + b) * d.K;
  double best = x[0];
  std::uint32_t arg = 0;
  for (std::uint32_t s = 1; s < d.K; ++s) {
    if (x[s] > best) {
      best = x[s];
      arg = s;
    }
  }
  log_weight[static_cast<std::uint64_t>(b) * d.roots + r] = best;
  assignment[static_cast<std::uint64_t>(b) * d.N + root] = arg;
}

// src/inference.cc:651-667
__global__ void MaxExpandRakes(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count, MaxBuffers buf,
                               std::uint32_t *assignment) {
  const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (tid >= static_cast<std::uint64_t>(count) * d.B) return;
  const std::uint32_t b = static_cast<std::uint32_t>(tid % d.B);
  const DeviceRake op = ops.rakes[offset + tid / d.B];
  std::uint32_t *a = assignment + static_cast<std::uint64_t>(b) * d.N;
  a[op.leaf] = buf.rake_choice[(static_cast<std::uint64_t>(op.branch) * d.B + b) * d.K + a[op.parent]];
}

__global__ void MaxExpandCompressions(DeviceOps ops, Dims d, std::uint32_t offset, std::uint32_t count,
                                      MaxBuffers buf, std::uint32_t *assignment) {
  const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (tid >= static_cast<std::uint64_t>(count) * d.B) return;
  const std::uint32_t b = static_cast<std::uint32_t>(tid % d.B);
  const DeviceCompress op = ops.compresses[offset + tid / d.B];
  std::uint32_t *a = assignment + static_cast<std::uint64_t>(b) * d.N;
  const std::uint64_t at = static_cast<std::uint64_t>(op.out - d.E) * d.B + b;
  a[op.middle] = buf.compress_choice[at * d.K * d.K + static_cast<std::uint64_t>(a[op.parent]) * d.K + a[op.child]];
}

// ------------------------------------------------------------------------
// Posterior sampling (FP64 scaled state)
// ------------------------------------------------------------------------

// Ascending inverse CDF over w / sum(w), skipping zero weights, falling back
// to the last supported state (src/inference.cc:128-146). Returns K when no
// state is supported.
template <class W> __device__ inline std::uint32_t InverseCdf(std::uint32_t K, W weight, double u) {
  double total = 0.0;
  for (std::uint32_t s = 0; s < K; ++s) total += weight(s);
  double cumulative = 0.0;
  std::uint32_t last = K;
  for (std::uint32_t s = 0; s < K; ++s) {
    const double w = weight(s);
    if (!(w > 0.0)) continue;
    last = s;
    cumulative += w / total;
    if (u < cumulative) return s;
  }
  return last;
}

template <typename U>
__global__ void SampleSeed(DeviceOps ops, Dims d, LinearBuffers buf, const U *uniforms, std::uint32_t *assignment,
                           std::uint32_t *failure) {
  const std::uint64_t tid = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (tid >= static_cast<std::uint64_t>(d.roots) * d.B) return;
  const std::uint32_t r = static_cast<std::uint32_t>(tid / d.B);
  const std::uint32_t b = static_cast<std::uint32_t>(tid % d.B);
  const std::uint32_t root = ops.roots[r];
  const double *x = static_cast<const double *>(buf.nodes) + (static_cast<std::uint64_t>(root) * d.B + b) * d.K;
  const std::uint64_t at = static_cast<std::uint64_t>(b) * d.N + root;
  const std::uint32_t s =
      InverseCdf(d.K, [&](std::uint32_t i) { return exp(x[i]); }, static_cast<double>(uniforms[at]));
  if (s >= d.K) *failure = 1;
  assignment[at] = s >= d.K ? 0 : s;
}

// The expansions draw each child's state given its parent's. G lanes take
// one (op, item): lane c forms the weight of state c with the same
// expression as a one-thread scan would, and the team's first lane draws
// with InverseCdf over the weights staged in LDS, so the draws do not depend
// on the team layout. Work per op is then K parallel exponentials and one
// short sequential scan.
template <typename U, int G>
__global__ __launch_bounds__(256) void SampleExpandRakes(DeviceOps ops, Dims d, std::uint32_t offset,
                                                         std::uint32_t count, LinearBuffers buf, const U *uniforms,
                                                         std::uint32_t *assignment, std::uint32_t *failure) {
  __shared__ double weights[256];
  const std::uint64_t item = (static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
  const std::uint32_t c = threadIdx.x % G;
  const bool active = item < static_cast<std::uint64_t>(count) * d.B;
  double w = 0.0;
  DeviceRake op{};
  std::uint32_t b = 0;
  if (active) {
    b = static_cast<std::uint32_t>(item % d.B);
    op = ops.rakes[offset + item / d.B];
    if (c < d.K) {
      const std::uint32_t *a = assignment + static_cast<std::uint64_t>(b) * d.N;
      const double row_c = SlotValue<double>(buf, d, op.slot, b, static_cast<std::uint64_t>(a[op.parent]) * d.K + c);
      const double *v = static_cast<const double *>(buf.nodes) + (static_cast<std::uint64_t>(op.leaf) * d.B + b) * d.K;
      w = row_c * exp(v[c]);
    }
  }
  weights[threadIdx.x] = w;
  __syncthreads();
  if (!active || c != 0) return;
  const double *team = weights + threadIdx.x;
  const std::uint64_t at = static_cast<std::uint64_t>(b) * d.N + op.leaf;
  const std::uint32_t s = InverseCdf(d.K, [&](std::uint32_t i) { return team[i]; }, static_cast<double>(uniforms[at]));
  if (s >= d.K) *failure = 1;
  assignment[at] = s >= d.K ? 0 : s;
}

// K = 8 and 16 with 16-byte aligned slots: K/4 lanes per op, one quad of
// columns each, so a wavefront holds four times as many ops (their chains
// of dependent loads: op, parent state, row) as with a lane per column. The
// weights are the same expression and the draw the same scan.
template <typename U, int K>
__global__ __launch_bounds__(256) void SampleExpandRakesQuad(DeviceOps ops, Dims d, std::uint32_t offset,
                                                             std::uint32_t count, LinearBuffers buf,
                                                             const U *uniforms, std::uint32_t *assignment,
                                                             std::uint32_t *failure) {
  constexpr int kT = K / 4;
  __shared__ double weights[256 * 4];
  const std::uint64_t item = (static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kT;
  const int q = static_cast<int>(threadIdx.x % kT);
  const bool active = item < static_cast<std::uint64_t>(count) * d.B;
  double *team = weights + (threadIdx.x - q) * 4;  // the op's K weights
  DeviceRake op{};
  std::uint32_t b = 0;
  if (active) {
    b = static_cast<std::uint32_t>(item % d.B);
    op = ops.rakes[offset + item / d.B];
    const std::uint32_t *a = assignment + static_cast<std::uint64_t>(b) * d.N;
    const Quad<double> row =
        LoadSlotQuad<double>(buf, d, op.slot, b, static_cast<std::uint64_t>(a[op.parent]) * K + 4 * q);
    const Quad<double> v =
        LoadQuad(static_cast<const double *>(buf.nodes) + (static_cast<std::uint64_t>(op.leaf) * d.B + b) * K + 4 * q);
#pragma unroll
    for (int r = 0; r < 4; ++r) team[4 * q + r] = row.v[r] * exp(v.v[r]);
  }
  __syncthreads();
  if (!active || q != 0) return;
  const std::uint64_t at = static_cast<std::uint64_t>(b) * d.N + op.leaf;
  const std::uint32_t s = InverseCdf(d.K, [&](std::uint32_t i) { return team[i]; }, static_cast<double>(uniforms[at]));
  if (s >= d.K) *failure = 1;
  assignment[at] = s >= d.K ? 0 : s;
}

template <typename U, int G>
__global__ __launch_bounds__(256) void SampleExpandCompressions(DeviceOps ops, Dims d, std::uint32_t offset,
                                                                std::uint32_t count, LinearBuffers buf,
                                                                const U *uniforms, std::uint32_t *assignment,
                                                                std::uint32_t *failure) {
  __shared__ double middle[256];   // mid[m] + r_m
  __shared__ double weights[256];
  const std::uint64_t item = (static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
  const std::uint32_t m = threadIdx.x % G;
  const bool active = item < static_cast<std::uint64_t>(count) * d.B;
  DeviceCompress op{};
  std::uint32_t b = 0;
  double x = kNegInf;
  if (active) {
    b = static_cast<std::uint32_t>(item % d.B);
    op = ops.compresses[offset + item / d.B];
    if (m < d.K)
      x = static_cast<const double *>(buf.nodes)[(static_cast<std::uint64_t>(op.middle) * d.B + b) * d.K + m] +
          RowScale(buf, d, op.right, b, m);
  }
  middle[threadIdx.x] = x;
  __syncthreads();
  const std::uint32_t base = threadIdx.x - m;
  double w = 0.0;
  if (active && m < d.K) {
    double top = kNegInf;
    for (std::uint32_t i = 0; i < d.K; ++i) top = fmax(top, middle[base + i]);
    if (top > kNegInf) {
      const std::uint32_t *a = assignment + static_cast<std::uint64_t>(b) * d.N;
      const double left = SlotValue<double>(buf, d, op.left, b, static_cast<std::uint64_t>(a[op.parent]) * d.K + m);
      const double right =
          SlotValue<double>(buf, d, op.right, b, static_cast<std::uint64_t>(m) * d.K + a[op.child]);
      w = left * exp(x - top) * right;
    }
  }
  weights[threadIdx.x] = w;
  __syncthreads();
  if (!active || m != 0) return;
  const double *team = weights + threadIdx.x;
  const std::uint64_t at = static_cast<std::uint64_t>(b) * d.N + op.middle;
  const std::uint32_t s = InverseCdf(d.K, [&](std::uint32_t i) { return team[i]; }, static_cast<double>(uniforms[at]));
  if (s >= d.K) *failure = 1;
  assignment[at] = s >= d.K ? 0 : s;
}

// The compression counterpart of SampleExpandRakesQuad: K/4 lanes per op,
// four middle states each; the same weights (left * exp(x - top) * right)
// and the same scan as SampleExpandCompressions.
template <typename U, int K>
__global__ __launch_bounds__(256) void SampleExpandCompressionsQuad(DeviceOps ops, Dims d, std::uint32_t offset,
                                                                    std::uint32_t count, LinearBuffers buf,
                                                                    const U *uniforms, std::uint32_t *assignment,
                                                                    std::uint32_t *failure) {
  constexpr int kT = K / 4;
  __shared__ double weights[256 * 4];
  const std::uint64_t item = (static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kT;
  const int q = static_cast<int>(threadIdx.x % kT);
  const bool active = item < static_cast<std::uint64_t>(count) * d.B;  // the same for a whole team
  double *team = weights + (threadIdx.x - q) * 4;
  DeviceCompress op{};
  std::uint32_t b = 0;
  if (active) {
    b = static_cast<std::uint32_t>(item % d.B);
    op = ops.compresses[offset + item / d.B];
    const Quad<double> mid = LoadQuad(static_cast<const double *>(buf.nodes) +
                                      (static_cast<std::uint64_t>(op.middle) * d.B + b) * K + 4 * q);
    double x[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) x[r] = mid.v[r] + RowScale(buf, d, op.right, b, 4 * q + r);
    double top = fmax(fmax(x[0], x[1]), fmax(x[2], x[3]));
#pragma unroll
    for (int o = 1; o < kT; o <<= 1) top = fmax(top, __shfl_xor(top, o, kT));
    double w[4] = {0.0, 0.0, 0.0, 0.0};
    if (top > kNegInf) {
      const std::uint32_t *a = assignment + static_cast<std::uint64_t>(b) * d.N;
      const Quad<double> left =
          LoadSlotQuad<double>(buf, d, op.left, b, static_cast<std::uint64_t>(a[op.parent]) * K + 4 * q);
      const std::uint32_t child = a[op.child];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double right = SlotValue<double>(buf, d, op.right, b, static_cast<std::uint64_t>(4 * q + r) * K + child);
        w[r] = left.v[r] * exp(x[r] - top) * right;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) team[4 * q + r] = w[r];
  }
  __syncthreads();
  if (!active || q != 0) return;
  const std::uint64_t at = static_cast<std::uint64_t>(b) * d.N + op.middle;
  const std::uint32_t s = InverseCdf(d.K, [&](std::uint32_t i) { return team[i]; }, static_cast<double>(uniforms[at]));
  if (s >= d.K) *failure = 1;
  assignment[at] = s >= d.K ? 0 : s;
}

// ------------------------------------------------------------------------
// Launch helpers
// ------------------------------------------------------------------------

constexpr unsigned kThreads = 256;

unsigned BlocksFor(std::uint64_t threads, unsigned block) {
  return static_cast<unsigned>((threads + block - 1) / block);
}

unsigned GroupBlock(int G) { return G <= 8 ? 256u : (G == 16 ? 128u : 64u); }

hipStream_t StreamOf(const Launch &l) { return static_cast<hipStream_t>(l.stream); }

template <typename T, int G>
void ForwardLevelT(const Launch &l, std::uint32_t primitive, std::uint32_t offset, std::uint32_t count,
                   const LinearBuffers &buf) {
  const unsigned block = GroupBlock(G);
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B * G, block);
  switch (primitive) {
    case 0:
      if (G == 4 && l.dims.K == 4)
        RakeForwardK4<T><<<BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B, 256), 256, 0, StreamOf(l)>>>(
            l.ops, l.dims, offset, count, buf);
      else
        RakeForward<T, G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      break;
    case 1:
      if (G == 4 && l.dims.K == 4)
        AddForwardK4<T, false><<<BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B, 256), 256, 0, StreamOf(l)>>>(
            l.ops, l.dims, offset, count, buf);
      else
        CombineForward<T, G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      break;
    case 2:
      if (G == 4 && l.dims.K == 4)
        AddForwardK4<T, true><<<BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B, 256), 256, 0, StreamOf(l)>>>(
            l.ops, l.dims, offset, count, buf);
      else
        AbsorbForward<T, G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      break;
    case 3: CompressForward<T, G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf); break;
  }
  Check(hipGetLastError(), "forward level launch");
}

template <typename T, int G>
void ReverseLevelT(const Launch &l, std::uint32_t primitive, std::uint32_t offset, std::uint32_t count,
                   const LinearBuffers &buf) {
  const unsigned block = GroupBlock(G);
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B * G, block);
  switch (primitive) {
    case 0:
      if (G == 4 && l.dims.K == 4)
        RakeReverseK4<T><<<BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B, 256), 256, 0, StreamOf(l)>>>(
            l.ops, l.dims, offset, count, buf);
      else
        RakeReverse<T, G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      break;
    case 1: CombineReverse<T, G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf); break;
    case 2: AbsorbReverse<T, G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf); break;
    case 3: {
      const std::size_t shared = static_cast<std::size_t>(block / G) * G * G * sizeof(double);
      CompressReverse<T, G><<<grid, block, shared, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      break;
    }
  }
  Check(hipGetLastError(), "reverse level launch");
}

template <typename U, int G>
void SampleExpandT(const Launch &l, std::uint32_t primitive, std::uint32_t offset, std::uint32_t count,
                   const LinearBuffers &buf, const U *uniforms, std::uint32_t *assignment, std::uint32_t *failure) {
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B * G, 256);
  const std::uint32_t K = l.dims.K;
  const bool quads = (K == 8 || K == 16 || K == 32) && (reinterpret_cast<std::uintptr_t>(buf.edges) & 15u) == 0;
  if (primitive == 0 && quads) {
    const unsigned quad_grid = BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B * (K / 4), 256);
    if (K == 8)
      SampleExpandRakesQuad<U, 8><<<quad_grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf, uniforms,
                                                                     assignment, failure);
    else if (K == 16)
      SampleExpandRakesQuad<U, 16><<<quad_grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf, uniforms,
                                                                      assignment, failure);
    else
      SampleExpandRakesQuad<U, 32><<<quad_grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf, uniforms,
                                                                      assignment, failure);
  } else if (primitive == 0) {
    SampleExpandRakes<U, G><<<grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf, uniforms, assignment,
                                                         failure);
  } else if (quads) {
    const unsigned quad_grid = BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B * (K / 4), 256);
    if (K == 8)
      SampleExpandCompressionsQuad<U, 8><<<quad_grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf,
                                                                            uniforms, assignment, failure);
    else if (K == 16)
      SampleExpandCompressionsQuad<U, 16><<<quad_grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf,
                                                                             uniforms, assignment, failure);
    else
      SampleExpandCompressionsQuad<U, 32><<<quad_grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf,
                                                                             uniforms, assignment, failure);
  } else
    SampleExpandCompressions<U, G><<<grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf, uniforms,
                                                                assignment, failure);
}

template <int G>
void MaxLevelT(const Launch &l, std::uint32_t primitive, std::uint32_t offset, std::uint32_t count,
               const MaxBuffers &buf) {
  // Team kernels: K*K/4 lanes per (op, item), 256 at K = 32, so a launch
  // there must stay below 2^31 work-items itself (the level chunks allow 64
  // lanes per item).
  const std::uint64_t lanes = static_cast<std::uint64_t>(count) * l.dims.B * l.dims.K * l.dims.K / 4;
  if ((l.dims.K == 8 || l.dims.K == 16 || l.dims.K == 32) && lanes < kMaxLaunchItems) {
    if (primitive == 0) {
      if (l.dims.K == 8)
        MaxRakeTeam<8><<<BlocksFor(lanes, 256), 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      else if (l.dims.K == 16)
        MaxRakeSteps<16><<<BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B * 16, 256), 256, 0, StreamOf(l)>>>(
            l.ops, l.dims, offset, count, buf);
      else
        MaxRakeSteps<32><<<BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B * 64, 256), 256, 0, StreamOf(l)>>>(
            l.ops, l.dims, offset, count, buf);
      Check(hipGetLastError(), "max-product level launch");
      return;
    }
    if (primitive == kMaxPiecePrimitive) {
      if (l.dims.K == 8)
        MaxPieceTeam<8><<<BlocksFor(lanes, 256), 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      else if (l.dims.K == 16)
        MaxPieceTeam<16><<<BlocksFor(lanes, 256), 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      else
        MaxPieceTeam<32><<<BlocksFor(lanes, 256), 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      Check(hipGetLastError(), "max-product level launch");
      return;
    }
    if (primitive == 3) {
      if (l.dims.K == 8)
        MaxCompressTeam<8><<<BlocksFor(lanes, 256), 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      else if (l.dims.K == 16)
        MaxCompressTeam<16><<<BlocksFor(lanes, 256), 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      else
        MaxCompressTeam<32><<<BlocksFor(lanes, 256), 256, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
      Check(hipGetLastError(), "max-product level launch");
      return;
    }
  }
  // Pieces are built only where the team kernel takes every piece level,
  // raw leaves only where it takes the first rake level.
  if (primitive == kMaxPiecePrimitive) {
    Check(hipErrorInvalidValue, "max-product pieces outside the team kernels");
    return;
  }
  if (primitive == 0 && (buf.raw_leaves != nullptr || buf.rake_direct != nullptr)) {
    Check(hipErrorInvalidValue, "raw leaf potentials or direct absorptions outside the team rake kernels");
    return;
  }
  const unsigned block = GroupBlock(G);
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B * G, block);
  switch (primitive) {
    case 0: MaxRake<G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf); break;
    case 1: MaxCombine<G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf); break;
    case 2: MaxAbsorb<G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf); break;
    case 3: MaxCompress<G><<<grid, block, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf); break;
  }
  Check(hipGetLastError(), "max-product level launch");
}

// Dispatches a runtime group width onto the template instantiations: runs
// CALL with the constant G bound.
#define THMM_WITH_GROUP(K, CALL)                    \
  switch (GroupWidth(K)) {                          \
    case 1: { constexpr int G = 1; CALL; } break;   \
    case 2: { constexpr int G = 2; CALL; } break;   \
    case 4: { constexpr int G = 4; CALL; } break;   \
    case 8: { constexpr int G = 8; CALL; } break;   \
    case 16: { constexpr int G = 16; CALL; } break; \
    case 32: { constexpr int G = 32; CALL; } break; \
    default: { constexpr int G = 64; CALL; } break; \
  }

template <typename In, typename Out, int kMode>
void LaunchTranspose(const Launch &l, const void *in, void *out, std::uint32_t X, std::uint32_t Y, double *scale,
                     const NodeList *only = nullptr) {
  constexpr Real in_real = sizeof(In) == 4 ? Real::kF32 : Real::kF64;
  constexpr Real out_real = sizeof(Out) == 4 ? Real::kF32 : Real::kF64;
  if (LaunchFastTranspose(l, kMode, in_real, out_real, in, out, X, Y, scale, only)) return;
  if (only != nullptr) {
    RecordLaunchError(hipErrorInvalidValue, "a partial node initialisation needs the team transpose");
    return;
  }
  const std::uint32_t ty = 64 / l.dims.K;
  const std::uint32_t x_tiles = (X + 63) / 64;
  const std::uint64_t tiles = static_cast<std::uint64_t>(x_tiles) * ((Y + ty - 1) / ty);
  if (tiles == 0) return;
  const unsigned grid = static_cast<unsigned>(tiles < (1u << 20) ? tiles : (1u << 20));
  TransposeVectors<In, Out, kMode><<<grid, kThreads, 0, StreamOf(l)>>>(
      static_cast<const In *>(in), static_cast<Out *>(out), X, Y, l.dims.K, scale, x_tiles);
}

}  // namespace

std::uint32_t GroupWidth(std::uint32_t states) {
  std::uint32_t g = 1;
  while (g < states) g *= 2;
  return g;
}

const char *LastLaunchError() {
  static thread_local std::string reported;
  reported = g_launch_error;
  g_launch_error.clear();
  return reported.c_str();
}

void LaunchInitNodes(const Launch &l, Real in_real, const void *input, void *nodes, double *node_scale,
                     const NodeList *only) {
  const Dims &d = l.dims;
  if (l.real == Real::kF32) {
    if (in_real == Real::kF32)
      LaunchTranspose<float, float, 1>(l, input, nodes, d.B, d.N, node_scale, only);
    else
      LaunchTranspose<double, float, 1>(l, input, nodes, d.B, d.N, node_scale, only);
  } else {
    if (in_real == Real::kF32)
      LaunchTranspose<float, double, 1>(l, input, nodes, d.B, d.N, node_scale, only);
    else
      LaunchTranspose<double, double, 1>(l, input, nodes, d.B, d.N, node_scale, only);
  }
  Check(hipGetLastError(), "node initialisation launch");
}

void LaunchInitNodesFromLogs(const Launch &l, const void *log_input, void *nodes, double *node_scale) {
  const Dims &d = l.dims;
  if (l.real == Real::kF32)
    LaunchTranspose<float, float, 2>(l, log_input, nodes, d.B, d.N, node_scale);
  else
    LaunchTranspose<double, double, 2>(l, log_input, nodes, d.B, d.N, node_scale);
  Check(hipGetLastError(), "log node initialisation launch");
}

void LaunchInitEdgesFromLogs(const Launch &l, const void *log_input, void *edges, double *edge_scale) {
  if (l.dims.E == 0) return;
  // Rows are K-vectors: the coalesced team kernel over E*K of them.
  const Real real = l.real;
  if (LaunchFastTranspose(l, 3, real, real, log_input, edges, 1, l.dims.E * l.dims.K, edge_scale)) {
    Check(hipGetLastError(), "log edge initialisation launch");
    return;
  }
  const unsigned grid = GridStrideBlocks(static_cast<std::uint64_t>(l.dims.E) * l.dims.K, kThreads);
  if (l.real == Real::kF32)
    InitEdgesFromLogs<float><<<grid, kThreads, 0, StreamOf(l)>>>(static_cast<const float *>(log_input),
                                                                static_cast<float *>(edges), edge_scale, l.dims);
  else
    InitEdgesFromLogs<double><<<grid, kThreads, 0, StreamOf(l)>>>(static_cast<const double *>(log_input),
                                                                 static_cast<double *>(edges), edge_scale, l.dims);
  Check(hipGetLastError(), "log edge initialisation launch");
}

void LaunchInitCategorical(const Launch &l, Real product_real, bool raw, const std::uint8_t *observations,
                           const std::uint32_t *observation_index_by_node, std::uint32_t observation_count,
                           std::uint32_t categories, const std::uint8_t *is_root, const void *root_potential,
                           const void *emissions, void *nodes, double *node_scale, const std::uint8_t *skip,
                           bool element_logs) {
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(l.dims.N) * l.dims.B, kThreads);
  const auto *obs = observations;
  const auto *index = observation_index_by_node;
  if (product_real == Real::kF32) {
    const auto *root = static_cast<const float *>(root_potential);
    const auto *em = static_cast<const float *>(emissions);
    if (raw)
      InitCategorical<float, float, false><<<grid, kThreads, 0, StreamOf(l)>>>(
          obs, index, observation_count, categories, is_root, root, em, static_cast<float *>(nodes), node_scale, l.dims, skip, element_logs);
    else if (l.real == Real::kF32)
      InitCategorical<float, float, true><<<grid, kThreads, 0, StreamOf(l)>>>(
          obs, index, observation_count, categories, is_root, root, em, static_cast<float *>(nodes), node_scale, l.dims, skip, element_logs);
    else
      InitCategorical<float, double, true><<<grid, kThreads, 0, StreamOf(l)>>>(
          obs, index, observation_count, categories, is_root, root, em, static_cast<double *>(nodes), node_scale,
          l.dims, skip, element_logs);
  } else {
    const auto *root = static_cast<const double *>(root_potential);
    const auto *em = static_cast<const double *>(emissions);
    if (raw)
      InitCategorical<double, double, false><<<grid, kThreads, 0, StreamOf(l)>>>(
          obs, index, observation_count, categories, is_root, root, em, static_cast<double *>(nodes), node_scale,
          l.dims, skip, element_logs);
    else
      InitCategorical<double, double, true><<<grid, kThreads, 0, StreamOf(l)>>>(
          obs, index, observation_count, categories, is_root, root, em, static_cast<double *>(nodes), node_scale,
          l.dims, skip, element_logs);
  }
  Check(hipGetLastError(), "categorical initialisation launch");
}

void LaunchSumOverBatch(const Launch &l, const double *log_z, double *sums) {
  if (l.dims.roots == 0) return;
  SumOverBatch<<<l.dims.roots, kSumThreads, 0, StreamOf(l)>>>(log_z, sums, l.dims);
  Check(hipGetLastError(), "batch sum launch");
}

// Whether a compression level goes to the team kernels: K = 4, 8, 16 or
// `max_k` (K*K/4 lanes per item, below 2^31 work-items) and 16-byte aligned
// slots and adjoint outputs.
bool TeamCompressionFits(const Launch &l, std::uint32_t ops, const LinearBuffers &buf, std::uint32_t max_k) {
  const std::uint32_t K = l.dims.K;
  const auto aligned = [](const void *p) { return (reinterpret_cast<std::uintptr_t>(p) & 15u) == 0; };
  return (K == 4 || K == 8 || K == 16 || K == max_k) &&
         static_cast<std::uint64_t>(ops) * l.dims.B * K * K / 4 < kMaxLaunchItems && aligned(buf.edges) &&
         aligned(buf.paths) && aligned(buf.path_adj) && aligned(buf.edge_out) && aligned(buf.node_adj);
}

template <typename T>
void LaunchCompressForwardTeam(const Launch &l, std::uint32_t first, std::uint32_t ops, const LinearBuffers &buf) {
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(ops) * l.dims.B * l.dims.K * l.dims.K / 4, 256);
  switch (l.dims.K) {
    case 4: CompressForwardTeam<T, 4><<<grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, first, ops, buf); break;
    case 8: CompressForwardTeam<T, 8><<<grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, first, ops, buf); break;
    case 16: CompressForwardTeam<T, 16><<<grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, first, ops, buf); break;
    default: CompressForwardTeam<T, 32><<<grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, first, ops, buf); break;
  }
  Check(hipGetLastError(), "compression level launch");
}

template <typename T>
void LaunchCompressReverseTeam(const Launch &l, std::uint32_t first, std::uint32_t ops, const LinearBuffers &buf) {
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(ops) * l.dims.B * l.dims.K * l.dims.K / 4, 256);
  switch (l.dims.K) {
    case 4: CompressReverseTeam<T, 4><<<grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, first, ops, buf); break;
    case 8: CompressReverseTeam<T, 8><<<grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, first, ops, buf); break;
    default: CompressReverseTeam<T, 16><<<grid, 256, 0, StreamOf(l)>>>(l.ops, l.dims, first, ops, buf); break;
  }
  Check(hipGetLastError(), "compression level launch");
}

void LaunchForwardLevel(const Launch &l, std::uint32_t primitive, std::uint32_t offset, std::uint32_t count,
                        const LinearBuffers &buf) {
  InOpChunks(offset, count, l.dims.B, [&](std::uint32_t first, std::uint32_t ops) {
    // Rakes with categorical leaves run on the generic kernels.
    const bool categorical = primitive == 0 && first < buf.cat.rakes;
    if (!categorical && LaunchFastLevel(l, false, primitive, first, ops, buf)) return;
    // Compressions at K = 4 and 8, and FP64 ones at K = 16 and 32: the team
    // kernel (16-byte aligned slots; K*K/4 lanes per item, below 2^31
    // work-items). FP32 compressions at K = 16 and 32 took the MFMA kernels.
    if (primitive == 3 && !buf.compress_log_edges && TeamCompressionFits(l, ops, buf, 32)) {
      if (l.real == Real::kF32)
        LaunchCompressForwardTeam<float>(l, first, ops, buf);
      else
        LaunchCompressForwardTeam<double>(l, first, ops, buf);
      return;
    }
    if (buf.edges_f32 && (primitive == 0 || primitive == 3)) {
      RecordLaunchError(hipErrorInvalidValue, "FP32 edges in an FP64 pass need the K = 8 and 16 level kernels");
      return;
    }
    if (buf.compress_log_edges && primitive == 3) {
      RecordLaunchError(hipErrorInvalidValue, "log edge compression inputs need the K = 32 and 64 MFMA kernels");
      return;
    }
    if (l.real == Real::kF32) {
      THMM_WITH_GROUP(l.dims.K, (ForwardLevelT<float, G>(l, primitive, first, ops, buf)));
    } else {
      THMM_WITH_GROUP(l.dims.K, (ForwardLevelT<double, G>(l, primitive, first, ops, buf)));
    }
  });
}

void LaunchFinish(const Launch &l, const LinearBuffers &buf, double *log_z) {
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(l.dims.roots) * l.dims.B, kThreads);
  if (l.real == Real::kF32)
    Finish<float><<<grid, kThreads, 0, StreamOf(l)>>>(l.ops, l.dims, buf, log_z);
  else
    Finish<double><<<grid, kThreads, 0, StreamOf(l)>>>(l.ops, l.dims, buf, log_z);
  Check(hipGetLastError(), "finish launch");
}

void LaunchSeedAdjoint(const Launch &l, const LinearBuffers &buf, const double *grad) {
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(l.dims.roots) * l.dims.B, kThreads);
  if (l.real == Real::kF32)
    SeedAdjoint<float><<<grid, kThreads, 0, StreamOf(l)>>>(l.ops, l.dims, buf, grad);
  else
    SeedAdjoint<double><<<grid, kThreads, 0, StreamOf(l)>>>(l.ops, l.dims, buf, grad);
  Check(hipGetLastError(), "adjoint seed launch");
}

void LaunchReverseLevel(const Launch &l, std::uint32_t primitive, std::uint32_t offset, std::uint32_t count,
                        const LinearBuffers &buf) {
  InOpChunks(offset, count, l.dims.B, [&](std::uint32_t first, std::uint32_t ops) {
    const bool categorical = primitive == 0 && first < buf.cat.rakes;
    if (!categorical && LaunchFastLevel(l, true, primitive, first, ops, buf)) return;
    if (buf.raw_leaves != nullptr) {
      RecordLaunchError(hipErrorInvalidValue, "raw leaf vectors need the team rake kernels");
      return;
    }
    // Compressions at K = 4, 8 and 16 (those the MFMA kernels left): the
    // team kernel.
    if (primitive == 3 && !buf.compress_log_edges && TeamCompressionFits(l, ops, buf, 16)) {
      if (l.real == Real::kF32)
        LaunchCompressReverseTeam<float>(l, first, ops, buf);
      else
        LaunchCompressReverseTeam<double>(l, first, ops, buf);
      return;
    }
    if (buf.compress_log_edges && primitive == 3) {
      RecordLaunchError(hipErrorInvalidValue, "log edge compression inputs need the K = 32 and 64 MFMA kernels");
      return;
    }
    if (l.real == Real::kF32) {
      THMM_WITH_GROUP(l.dims.K, (ReverseLevelT<float, G>(l, primitive, first, ops, buf)));
    } else {
      THMM_WITH_GROUP(l.dims.K, (ReverseLevelT<double, G>(l, primitive, first, ops, buf)));
    }
  });
}

void LaunchNodeAdjointOut(const Launch &l, const void *node_adj, void *out) {
  const Dims &d = l.dims;
  if (l.real == Real::kF32)
    LaunchTranspose<float, float, 0>(l, node_adj, out, d.N, d.B, nullptr);
  else
    LaunchTranspose<double, double, 0>(l, node_adj, out, d.N, d.B, nullptr);
  Check(hipGetLastError(), "node marginal transpose launch");
}

namespace {

// One thread per (e, p, c): the batch-summed adjoint times the rate of change
// of log P_e[p][c] with respect to t_e.
// Batch-summed edge adjoint M_e[p][c] times d log P_e[p][c] / d t_e. A block
// covers 64 consecutive entries (e, p, c); its four wavefronts sum batch
// items b = w, w + 4, ... of them (each load instruction reads 64
// consecutive values), and the four partial masses are added in LDS in a
// fixed order.
template <typename T>
__global__ __launch_bounds__(256) void BranchChainTerms(const T *edge_out, const T *edges, const double *rate,
                                                        double *terms, Dims d) {
  __shared__ double part[4][64];
  const std::uint64_t kk = static_cast<std::uint64_t>(d.K) * d.K;
  const std::uint64_t total = static_cast<std::uint64_t>(d.E) * kk;
  const std::uint32_t lane = threadIdx.x % 64, slice = threadIdx.x / 64;
  for (std::uint64_t base = static_cast<std::uint64_t>(blockIdx.x) * 64; base < total;
       base += static_cast<std::uint64_t>(gridDim.x) * 64) {
    const std::uint64_t tid = base + lane;
    double mass = 0.0;
    if (tid < total) {
#pragma unroll 8
      for (std::uint32_t b = slice; b < d.B; b += 4) mass += static_cast<double>(edge_out[b * total + tid]);
    }
    part[slice][lane] = mass;
    __syncthreads();
    if (slice == 0 && tid < total) {
      mass = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
      const std::uint64_t e = tid / kk;
      const std::uint32_t p = static_cast<std::uint32_t>((tid % kk) / d.K), c = static_cast<std::uint32_t>(tid % d.K);
      const T *P = edges + e * kk;
      const double value = static_cast<double>(P[p * d.K + c]);
      double term = 0.0;
      if (value > 0.0 && mass != 0.0) {
        double slope = 0.0;
        for (std::uint32_t j = 0; j < d.K; ++j) slope += rate[p * d.K + j] * static_cast<double>(P[j * d.K + c]);
        term = mass * slope / value;
      }
      terms[tid] = term;
    }
    __syncthreads();
  }
}

// in [rows][cols] bytes -> out [cols][rows] through a 64 x 64 LDS tile; the
// grid covers the tiles (block i: column tile i % col_tiles, row tile
// i / col_tiles, stepping by the grid).
__global__ __launch_bounds__(256) void TransposeBytes(const std::uint8_t *in, std::uint8_t *out, std::uint32_t rows,
                                                      std::uint32_t cols, std::uint64_t tiles,
                                                      std::uint32_t col_tiles) {
  __shared__ std::uint8_t tile[64][65];
  for (std::uint64_t i = blockIdx.x; i < tiles; i += gridDim.x) {
    const std::uint64_t r0 = (i / col_tiles) * 64, c0 = (i % col_tiles) * 64;
    for (std::uint32_t k = threadIdx.x; k < 64 * 64; k += blockDim.x) {
      const std::uint64_t r = r0 + k / 64, c = c0 + k % 64;
      if (r < rows && c < cols) tile[k / 64][k % 64] = in[r * cols + c];
    }
    __syncthreads();
    for (std::uint32_t k = threadIdx.x; k < 64 * 64; k += blockDim.x) {
      const std::uint64_t c = c0 + k / 64, r = r0 + k % 64;
      if (r < rows && c < cols) out[c * rows + r] = tile[k % 64][k / 64];
    }
    __syncthreads();
  }
}

__global__ void BranchChainSums(const double *terms, double *grad, Dims d) {
  const std::uint64_t e = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= d.E) return;
  const std::uint64_t kk = static_cast<std::uint64_t>(d.K) * d.K;
  double total = 0.0;
  for (std::uint64_t i = 0; i < kk; ++i) total += terms[e * kk + i];
  grad[e] = total;
}

}  // namespace

namespace {

// One thread per (e, i, j); exp(lambda_k t_e) is recomputed per entry (K of
// them), which is negligible next to the K-term sum at the state counts of
// substitution models.
template <typename T>
__global__ void ExpmEdges(const double *V, const double *lambda, const double *V_inv, const double *t, T *edges,
                          Dims d) {
  const std::uint64_t kk = static_cast<std::uint64_t>(d.K) * d.K;
  for (std::uint64_t tid = GridThread(); tid < kk * d.E; tid += GridStride()) {
    const std::uint64_t e = tid / kk;
    const std::uint32_t i = static_cast<std::uint32_t>(tid % kk) / d.K, j = static_cast<std::uint32_t>(tid % kk) % d.K;
    double sum = 0.0;
    for (std::uint32_t k = 0; k < d.K; ++k) sum += V[i * d.K + k] * exp(lambda[k] * t[e]) * V_inv[k * d.K + j];
    edges[tid] = static_cast<T>(fmax(sum, 0.0));
  }
}

}  // namespace

void LaunchExpmEdges(const Launch &l, const double *V, const double *lambda, const double *V_inv,
                     const double *branch_lengths, void *edges) {
  const std::uint64_t entries = static_cast<std::uint64_t>(l.dims.E) * l.dims.K * l.dims.K;
  if (entries == 0) return;
  if (l.real == Real::kF32)
    ExpmEdges<float><<<GridStrideBlocks(entries, kThreads), kThreads, 0, StreamOf(l)>>>(
        V, lambda, V_inv, branch_lengths, static_cast<float *>(edges), l.dims);
  else
    ExpmEdges<double><<<GridStrideBlocks(entries, kThreads), kThreads, 0, StreamOf(l)>>>(
        V, lambda, V_inv, branch_lengths, static_cast<double *>(edges), l.dims);
  Check(hipGetLastError(), "expm edge launch");
}

void LaunchBranchLengthChain(const Launch &l, const void *edge_out, const void *edges, const double *rate,
                             double *terms, double *grad) {
  const std::uint64_t entries = static_cast<std::uint64_t>(l.dims.E) * l.dims.K * l.dims.K;
  if (entries == 0) return;
  const unsigned blocks = GridStrideBlocks(entries, 64);
  if (l.real == Real::kF32)
    BranchChainTerms<float><<<blocks, 256, 0, StreamOf(l)>>>(static_cast<const float *>(edge_out),
                                                             static_cast<const float *>(edges), rate, terms, l.dims);
  else
    BranchChainTerms<double><<<blocks, 256, 0, StreamOf(l)>>>(static_cast<const double *>(edge_out),
                                                              static_cast<const double *>(edges), rate, terms, l.dims);
  BranchChainSums<<<BlocksFor(l.dims.E, kThreads), kThreads, 0, StreamOf(l)>>>(terms, grad, l.dims);
  Check(hipGetLastError(), "branch-length chain launch");
}

void LaunchTransposeObservations(const Launch &l, const std::uint8_t *in, std::uint8_t *out, std::uint32_t rows,
                                 std::uint32_t cols) {
  if (rows == 0 || cols == 0) return;
  const std::uint32_t col_tiles = (cols + 63) / 64;
  const std::uint64_t tiles = static_cast<std::uint64_t>(col_tiles) * ((rows + 63) / 64);
  const unsigned grid = static_cast<unsigned>(tiles < (1u << 20) ? tiles : (1u << 20));
  TransposeBytes<<<grid, 256, 0, StreamOf(l)>>>(in, out, rows, cols, tiles, col_tiles);
  Check(hipGetLastError(), "observation transpose launch");
}

void LaunchReduceEdgesOverBatch(const Launch &l, const void *edge_out, void *out) {
  const std::uint64_t total = static_cast<std::uint64_t>(l.dims.E) * l.dims.K * l.dims.K;
  if (total == 0) return;
  const unsigned grid = GridStrideBlocks(total, kThreads);
  if (l.real == Real::kF32)
    ReduceEdges<float><<<grid, kThreads, 0, StreamOf(l)>>>(static_cast<const float *>(edge_out),
                                                          static_cast<float *>(out), l.dims);
  else
    ReduceEdges<double><<<grid, kThreads, 0, StreamOf(l)>>>(static_cast<const double *>(edge_out),
                                                           static_cast<double *>(out), l.dims);
  Check(hipGetLastError(), "edge reduction launch");
}

void LaunchWidenF32(const Launch &l, const float *in, double *out, std::size_t count) {
  if (count == 0) return;
  Convert<float, double><<<GridStrideBlocks(count, kThreads), kThreads, 0, StreamOf(l)>>>(in, out, count);
  Check(hipGetLastError(), "widen launch");
}

void LaunchCastF64(const Launch &l, const double *in, void *out, std::size_t count) {
  if (count == 0) return;
  const unsigned grid = GridStrideBlocks(count, kThreads);
  if (l.real == Real::kF32)
    Convert<double, float><<<grid, kThreads, 0, StreamOf(l)>>>(in, static_cast<float *>(out), count);
  else
    Convert<double, double><<<grid, kThreads, 0, StreamOf(l)>>>(in, static_cast<double *>(out), count);
  Check(hipGetLastError(), "cast launch");
}

void LaunchMaxInit(const Launch &l, Real in_real, const void *nodes_in, const void *edges_in, double *nodes,
                   double *edges, const NodeList *only) {
  const std::uint32_t *ids = only != nullptr ? only->ids : nullptr;
  const std::uint32_t listed = only != nullptr ? only->count : 0;
  const std::uint64_t total = static_cast<std::uint64_t>(only != nullptr ? listed : l.dims.N) * l.dims.B * l.dims.K +
                              static_cast<std::uint64_t>(l.dims.E) * l.dims.K * l.dims.K;
  if (total == 0) return;
  const auto aligned = [](const void *p) { return (reinterpret_cast<std::uintptr_t>(p) & 15u) == 0; };
  if (l.dims.K % 4 == 0 && aligned(nodes_in) && aligned(nodes) && (l.dims.E == 0 || (aligned(edges_in) && aligned(edges)))) {
    const unsigned grid = GridStrideBlocks(total / 4, kThreads);
    if (in_real == Real::kF32)
      MaxInitQuad<float><<<grid, kThreads, 0, StreamOf(l)>>>(static_cast<const float *>(nodes_in),
                                                            static_cast<const float *>(edges_in), nodes, edges, l.dims,
                                                            ids, listed);
    else
      MaxInitQuad<double><<<grid, kThreads, 0, StreamOf(l)>>>(static_cast<const double *>(nodes_in),
                                                             static_cast<const double *>(edges_in), nodes, edges,
                                                             l.dims, ids, listed);
    Check(hipGetLastError(), "max-product initialisation launch");
    return;
  }
  const unsigned grid = GridStrideBlocks(total, kThreads);
  if (in_real == Real::kF32)
    MaxInit<float><<<grid, kThreads, 0, StreamOf(l)>>>(static_cast<const float *>(nodes_in),
                                                      static_cast<const float *>(edges_in), nodes, edges, l.dims,
                                                      ids, listed);
  else
    MaxInit<double><<<grid, kThreads, 0, StreamOf(l)>>>(static_cast<const double *>(nodes_in),
                                                       static_cast<const double *>(edges_in), nodes, edges, l.dims,
                                                       ids, listed);
  Check(hipGetLastError(), "max-product initialisation launch");
}

void LaunchMaxLevel(const Launch &l, std::uint32_t primitive, std::uint32_t offset, std::uint32_t count,
                    const MaxBuffers &buf) {
  InOpChunks(offset, count, l.dims.B, [&](std::uint32_t first, std::uint32_t ops) {
    THMM_WITH_GROUP(l.dims.K, (MaxLevelT<G>(l, primitive, first, ops, buf)));
  });
}

void LaunchMaxFinish(const Launch &l, const MaxBuffers &buf, double *log_weight, std::uint32_t *assignment) {
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(l.dims.roots) * l.dims.B, kThreads);
  MaxFinish<<<grid, kThreads, 0, StreamOf(l)>>>(l.ops, l.dims, buf, log_weight, assignment);
  Check(hipGetLastError(), "max-product finish launch");
}

void LaunchMaxExpandLevel(const Launch &l, std::uint32_t primitive, std::uint32_t offset, std::uint32_t count,
                          const MaxBuffers &buf, std::uint32_t *assignment) {
  if (count == 0) return;
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(count) * l.dims.B, kThreads);
  if (primitive == 0)
    MaxExpandRakes<<<grid, kThreads, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf, assignment);
  else if (primitive == 3)
    MaxExpandCompressions<<<grid, kThreads, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf, assignment);
  Check(hipGetLastError(), "MAP expansion launch");
}

namespace {
__global__ void ClearWord(std::uint32_t *word) { *word = 0; }
}  // namespace

void LaunchClearWord(const Launch &l, std::uint32_t *word) {
  ClearWord<<<1, 1, 0, StreamOf(l)>>>(word);
  Check(hipGetLastError(), "flag clear launch");
}

void LaunchSampleSeed(const Launch &l, const LinearBuffers &buf, Real in_real, const void *uniforms,
                      std::uint32_t *assignment, std::uint32_t *failure) {
  const unsigned grid = BlocksFor(static_cast<std::uint64_t>(l.dims.roots) * l.dims.B, kThreads);
  if (in_real == Real::kF32)
    SampleSeed<float><<<grid, kThreads, 0, StreamOf(l)>>>(l.ops, l.dims, buf, static_cast<const float *>(uniforms),
                                                         assignment, failure);
  else
    SampleSeed<double><<<grid, kThreads, 0, StreamOf(l)>>>(l.ops, l.dims, buf, static_cast<const double *>(uniforms),
                                                          assignment, failure);
  Check(hipGetLastError(), "sample seed launch");
}

void LaunchSampleExpandLevel(const Launch &l, std::uint32_t primitive, std::uint32_t offset, std::uint32_t count,
                             const LinearBuffers &buf, Real in_real, const void *uniforms, std::uint32_t *assignment,
                             std::uint32_t *failure) {
  if (count == 0 || (primitive != 0 && primitive != 3)) return;
  InOpChunks(offset, count, l.dims.B, [&](std::uint32_t first, std::uint32_t ops) {
    if (in_real == Real::kF32) {
      THMM_WITH_GROUP(l.dims.K, (SampleExpandT<float, G>(l, primitive, first, ops, buf,
                                                         static_cast<const float *>(uniforms), assignment, failure)));
    } else {
      THMM_WITH_GROUP(l.dims.K, (SampleExpandT<double, G>(l, primitive, first, ops, buf,
                                                          static_cast<const double *>(uniforms), assignment, failure)));
    }
  });
  Check(hipGetLastError(), "sample expansion launch");
}

}  // namespace thmm
