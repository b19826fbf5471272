// Coalesced and MFMA kernels for the hot levels (see kernels.hip for the
// number representation and the generic forms these replace).
//
// Team layout (rakes, K in {8, 16, 32}): a team of T lanes owns one item.
// The item's K x K matrix is read and written as row-major 16-byte quads:
// in step j lane l moves quad j*T + l, i.e. row j*RPS + l/LPR, columns
// (l%LPR)*4 .. +3, so every memory instruction of a team touches one
// contiguous T*16-byte span (LPR = K/4 lanes per row, RPS = 4T/K rows per
// step). Row sums reduce over the LPR lanes of a row, column sums over the
// lanes that share columns.
//
// MFMA compression (FP32, K = 16): one wavefront per item; L and R are
// staged through LDS with coalesced 16-byte loads and the 16 x 16 x 16
// products run on v_mfma_f32_16x16x4_f32 (four k-steps). Operand layouts:
// A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in register i%4
// of lane j + 16(i/4). Every MFMA operand lies in [0, 1]: L and R are
// row-normalised, the middle weights are exp(w - max w), and in the reverse
// pass the ratio matrix adj/S is row-normalised with its row scales (and
// their maximum) kept in FP64.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.h"
#include "launch.h"

namespace thmm {
namespace {

using namespace dev;

template <int K> struct Team {
  static constexpr int kLanesPerRow = K / 4;
  static constexpr int kLanes = K <= 4 ? 4 : (K <= 16 ? 16 : 64);
  static constexpr int kRowsPerStep = 4 * kLanes / K;
  static constexpr int kSteps = K / kRowsPerStep;
};
static_assert(Team<8>::kSteps == 1 && Team<16>::kSteps == 4 && Team<32>::kSteps == 4, "team tiling");
// One logarithm per lane in RakeForwardTeam needs at most one step per lane
// of a row.
static_assert(Team<8>::kSteps <= Team<8>::kLanesPerRow && Team<16>::kSteps <= Team<16>::kLanesPerRow &&
                  Team<32>::kSteps <= Team<32>::kLanesPerRow,
              "row ownership");

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
#pragma unroll
  for (int i = 0; i < 4; ++i) e[i] = Exp(vq.v[i]);
  const T *P = Slot<T>(buf, d, op.slot, b);
  // After the row reduction all kLpr lanes of a row hold its sum; the lane
  // with l % kLpr == j % kLpr takes the logarithm, so each row's log is
  // formed once and the team's rows are stored by one contiguous store.
  T sums[kJ];
#pragma unroll
  for (int j = 0; j < kJ; ++j) {
    const Quad<T> p = LoadQuad(P + static_cast<std::uint32_t>(j * kT + l) * 4);
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
  const double message =
      row >= 0 && part > T(0) ? RowScale(buf, d, op.slot, b, row) + log(static_cast<double>(part)) : kNegInf;
  double top = message;
#pragma unroll
  for (int o = kT / 2; o > 0; o >>= 1) top = fmax(top, __shfl_xor(top, o, kT));
  const std::uint64_t at = static_cast<std::uint64_t>(op.branch) * d.B + b;
  const bool live = top > kNegInf;
  if (row >= 0) static_cast<T *>(buf.branches)[at * K + row] = live ? static_cast<T>(message - top) : NegInf<T>();
  if (l == 0) buf.branch_scale[at] = live ? buf.node_scale[leaf_at] + top : kNegInf;
}

// Rake adjoint (src/inference.cc:488-518): with t = P[p][c] exp(v[c]) and
// m[p] = sum_c t, adj_P = adj_m[p] t / m[p] (ratio in FP64) and
// adj_v[c] = sum_p adj_P[p][c].
template <typename T, int K>
__global__ __launch_bounds__(256) void RakeReverseTeam(DeviceOps ops, Dims d, std::uint32_t offset,
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
#pragma unroll
  for (int i = 0; i < 4; ++i) e[i] = Exp(vq.v[i]);
  const T *P = Slot<T>(buf, d, op.slot, b);
  T *adj_p = SlotAdjoint<T>(buf, d, op.slot, b);
  const T *adj_m = static_cast<const T *>(buf.branch_adj) + (static_cast<std::uint64_t>(op.branch) * d.B + b) * K;
  double column[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < kJ; ++j) {
    const std::uint32_t quad = static_cast<std::uint32_t>(j * kT + l) * 4;
    const Quad<T> p = LoadQuad(P + quad);
    T t[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = p.v[i] * e[i];
    T part = t[0] + t[1] + t[2] + t[3];
#pragma unroll
    for (int o = kLpr / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, kT);
    const T a = adj_m[j * kRps + l / kLpr];
    const double g = (a != T(0) && part > T(0)) ? static_cast<double>(a) / static_cast<double>(part) : 0.0;
    Quad<T> out;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double c = g * static_cast<double>(t[i]);
      out.v[i] = static_cast<T>(c);
      column[i] += c;
    }
    StoreQuad(adj_p + quad, out);
  }
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

// ---------------------------------------------------------------------------
// FP32 MFMA compression, K = 16
// ---------------------------------------------------------------------------

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kPad = 17;  // LDS row stride (floats)
constexpr int kWaves = 4;  // items (wavefronts) per block

struct CompressTiles {
  float L[16 * kPad];
  float R[16 * kPad];
  float G[16 * kPad];
  double gscale[16];
};

// Stages L and R of an item in LDS and returns the lane's middle weight
// e[m] for m = lane % 16, with W = max_m (mid[m] + r_m) through *top.
__device__ inline float StageCompression(const LinearBuffers &buf, const Dims &d, const DeviceCompress &op,
                                         std::uint32_t b, int lane, CompressTiles &tiles, double *top) {
  const float *L = Slot<float>(buf, d, op.left, b);
  const float *R = Slot<float>(buf, d, op.right, b);
  const Quad<float> ql = LoadQuad(L + lane * 4);
  const Quad<float> qr = LoadQuad(R + lane * 4);
  const int row = lane / 4, col = (lane % 4) * 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    tiles.L[row * kPad + col + i] = ql.v[i];
    tiles.R[row * kPad + col + i] = qr.v[i];
  }
  const int m = lane % 16;
  const std::uint64_t mid_at = static_cast<std::uint64_t>(op.middle) * d.B + b;
  const double w = static_cast<double>(static_cast<const float *>(buf.nodes)[mid_at * 16 + m]) +
                   RowScale(buf, d, op.right, b, m);
  double w_top = w;
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) w_top = fmax(w_top, __shfl_xor(w_top, o, 16));
  *top = w_top;
  return w_top > kNegInf ? expf(static_cast<float>(w - w_top)) : 0.0f;
}

// S = L diag(e) R on MFMA; lane holds S[4(lane/16) + i][lane % 16].
__device__ inline floatx4 MiddleProduct(const CompressTiles &tiles, float e_mine, int lane) {
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const int k = 4 * kb + lane / 16;
    const float ek = __shfl(e_mine, k, 64);
    const float a = tiles.L[(lane % 16) * kPad + k] * ek;
    const float bv = tiles.R[k * kPad + lane % 16];
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv, acc, 0, 0, 0);
  }
  return acc;
}

__global__ __launch_bounds__(256) void CompressForwardMfma16(DeviceOps ops, Dims d, std::uint32_t offset,
                                                             std::uint32_t count, LinearBuffers buf) {
  __shared__ CompressTiles shared[kWaves];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const std::uint32_t item = blockIdx.x * kWaves + wave;
  if (item >= count * d.B) return;  // whole wavefront; no block barrier below
  CompressTiles &tiles = shared[wave];
  const std::uint32_t b = item % d.B;
  const DeviceCompress op = ops.compresses[offset + item / d.B];
  double w_top = 0.0;
  const float e_mine = StageCompression(buf, d, op, b, lane, tiles, &w_top);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const floatx4 s = MiddleProduct(tiles, e_mine, lane);
  const std::uint64_t mid_at = static_cast<std::uint64_t>(op.middle) * d.B + b;
  const std::uint64_t out_at = static_cast<std::uint64_t>(op.out - d.E) * d.B + b;
  float *out = static_cast<float *>(buf.paths) + out_at * 256;
  const int c = lane % 16;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int p = 4 * (lane / 16) + i;
    float rmax = s[i];
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) rmax = fmaxf(rmax, __shfl_xor(rmax, o, 16));
    const bool live = rmax > 0.0f && w_top > kNegInf;
    out[p * 16 + c] = live ? s[i] / rmax : 0.0f;
    if (c == 0)
      buf.path_scale[out_at * 16 + p] =
          live ? RowScale(buf, d, op.left, b, p) + buf.node_scale[mid_at] + w_top + log(static_cast<double>(rmax))
               : kNegInf;
  }
}

// Compression adjoint on MFMA (src/inference.cc:416-464). With S = L' R,
// L' = L diag(e), Gr = adj_out / S row-normalised as Gr[p][c] = g_p Gh[p][c]
// and Gamma = max_p g_p:
//   adj_L[p][m] = L'[p][m] g_p (Gh R^T)[p][m]
//   adj_R[m][c] = R[m][c] e[m] Gamma (A^T Gh)[m][c],  A[p][m] = L[p][m] g_p / Gamma
//   adj_mid[m]  = sum_c adj_R[m][c]
__global__ __launch_bounds__(256) void CompressReverseMfma16(DeviceOps ops, Dims d, std::uint32_t offset,
                                                             std::uint32_t count, LinearBuffers buf) {
  __shared__ CompressTiles shared[kWaves];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const std::uint32_t item = blockIdx.x * kWaves + wave;
  if (item >= count * d.B) return;
  CompressTiles &tiles = shared[wave];
  const std::uint32_t b = item % d.B;
  const DeviceCompress op = ops.compresses[offset + item / d.B];
  double w_top = 0.0;
  const float e_mine = StageCompression(buf, d, op, b, lane, tiles, &w_top);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const floatx4 s = MiddleProduct(tiles, e_mine, lane);
  const std::uint64_t out_at = static_cast<std::uint64_t>(op.out - d.E) * d.B + b;
  const float *adj_out = static_cast<const float *>(buf.path_adj) + out_at * 256;
  const int c = lane % 16;
  // Row-normalised ratios Gh and their FP64 row scales g_p.
  double gamma = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int p = 4 * (lane / 16) + i;
    const float a = adj_out[p * 16 + c];
    const double gr = (a != 0.0f && s[i] > 0.0f) ? static_cast<double>(a) / static_cast<double>(s[i]) : 0.0;
    double g = gr;
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) g = fmax(g, __shfl_xor(g, o, 16));
    tiles.G[p * kPad + c] = g > 0.0 ? static_cast<float>(gr / g) : 0.0f;
    if (c == 0) tiles.gscale[p] = g;
    gamma = fmax(gamma, g);
  }
#pragma unroll
  for (int o = 16; o < 64; o <<= 1) gamma = fmax(gamma, __shfl_xor(gamma, o, 64));
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  // H = Gh R^T; lane holds H[p = 4(lane/16) + i][m = lane % 16].
  floatx4 h = {0.f, 0.f, 0.f, 0.f};
  // Q = A^T Gh; lane holds Q[m = 4(lane/16) + i][c = lane % 16].
  floatx4 q = {0.f, 0.f, 0.f, 0.f};
  const double inverse_gamma = gamma > 0.0 ? 1.0 / gamma : 0.0;
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const int k = 4 * kb + lane / 16;
    const int r = lane % 16;
    h = __builtin_amdgcn_mfma_f32_16x16x4f32(tiles.G[r * kPad + k], tiles.R[r * kPad + k], h, 0, 0, 0);
    const float a = tiles.L[k * kPad + r] * static_cast<float>(tiles.gscale[k] * inverse_gamma);
    q = __builtin_amdgcn_mfma_f32_16x16x4f32(a, tiles.G[k * kPad + r], q, 0, 0, 0);
  }
  float *adj_left = SlotAdjoint<float>(buf, d, op.left, b);
  float *adj_right = SlotAdjoint<float>(buf, d, op.right, b);
  const std::uint64_t mid_at = static_cast<std::uint64_t>(op.middle) * d.B + b;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = 4 * (lane / 16) + i;
    // adj_L[p = row][m = c]
    const double left = static_cast<double>(tiles.L[row * kPad + c] * e_mine) * tiles.gscale[row] *
                        static_cast<double>(h[i]);
    adj_left[row * 16 + c] = static_cast<float>(left);
    // adj_R[m = row][c]
    const float e_row = __shfl(e_mine, row, 64);
    const double right = static_cast<double>(tiles.R[row * kPad + c]) * static_cast<double>(e_row) * gamma *
                         static_cast<double>(q[i]);
    adj_right[row * 16 + c] = static_cast<float>(right);
    double total = right;
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) total += __shfl_xor(total, o, 16);
    if (c == 0) static_cast<float *>(buf.node_adj)[mid_at * 16 + row] = static_cast<float>(total);
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
// goes to `scale` in FP64 (-inf for a vector without support).
template <typename T> __device__ inline T LogOf(T v) { return v > T(0) ? log(v) : NegInf<T>(); }
template <> __device__ inline float LogOf(float v) { return v > 0.0f ? logf(v) : NegInf<float>(); }

template <typename In, typename Out, int kMode, int K>
__global__ __launch_bounds__(256) void TransposeTeam(const In *in, Out *out, std::uint32_t X, std::uint32_t Y,
                                                     double *scale) {
  constexpr int kLanes = K / 4;
  const std::uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const std::uint32_t v = tid / kLanes;
  const int l = static_cast<int>(tid % kLanes);
  if (v >= X * Y) return;
  const std::uint32_t x = v / Y, y = v % Y;
  const std::uint64_t to = static_cast<std::uint64_t>(y) * X + x;
  const Quad<In> q = LoadQuad(in + static_cast<std::uint64_t>(v) * K + l * 4);
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
    for (int i = 0; i < 4; ++i) o.v[i] = live ? static_cast<Out>(w[i] - top) : NegInf<Out>();
    if (l == 0) scale[to] = live ? static_cast<double>(top) : kNegInf;
  }
  StoreQuad(out + to * K + l * 4, o);
}

constexpr unsigned kBlock = 256;

bool Aligned16(const void *p) { return (reinterpret_cast<std::uintptr_t>(p) & 15u) == 0; }

unsigned Blocks(std::uint64_t threads) { return static_cast<unsigned>((threads + kBlock - 1) / kBlock); }

hipStream_t StreamOf(const Launch &l) { return static_cast<hipStream_t>(l.stream); }

template <typename T, int K> void LaunchRakeTeam(const Launch &l, bool reverse, std::uint32_t offset, std::uint32_t count,
                                                 const LinearBuffers &buf) {
  const unsigned grid = Blocks(static_cast<std::uint64_t>(count) * l.dims.B * Team<K>::kLanes);
  if (reverse)
    RakeReverseTeam<T, K><<<grid, kBlock, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
  else
    RakeForwardTeam<T, K><<<grid, kBlock, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
}

}  // namespace

bool LaunchFastLevel(const Launch &l, bool reverse, std::uint32_t primitive, std::uint32_t offset,
                     std::uint32_t count, const LinearBuffers &buf) {
  const std::uint32_t K = l.dims.K;
  // 32-bit item indexing: the grid must address every (op, item, lane).
  // Caller-owned edge inputs and outputs must allow 16-byte accesses.
  if (!Aligned16(buf.edges) || !Aligned16(buf.edge_out) || !Aligned16(buf.node_adj)) return false;
  if (static_cast<std::uint64_t>(count) * l.dims.B * 64 >= (1ull << 32)) return false;
  if (primitive == 0 && (K == 8 || K == 16 || K == 32)) {
    if (l.real == Real::kF32) {
      if (K == 8) LaunchRakeTeam<float, 8>(l, reverse, offset, count, buf);
      if (K == 16) LaunchRakeTeam<float, 16>(l, reverse, offset, count, buf);
      if (K == 32) LaunchRakeTeam<float, 32>(l, reverse, offset, count, buf);
    } else {
      if (K == 8) LaunchRakeTeam<double, 8>(l, reverse, offset, count, buf);
      if (K == 16) LaunchRakeTeam<double, 16>(l, reverse, offset, count, buf);
      if (K == 32) LaunchRakeTeam<double, 32>(l, reverse, offset, count, buf);
    }
    RecordLaunchError(hipGetLastError(), "rake level launch");
    return true;
  }
  if (primitive == 3 && K == 16 && l.real == Real::kF32) {
    const unsigned grid = static_cast<unsigned>((static_cast<std::uint64_t>(count) * l.dims.B + kWaves - 1) / kWaves);
    if (reverse)
      CompressReverseMfma16<<<grid, kBlock, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
    else
      CompressForwardMfma16<<<grid, kBlock, 0, StreamOf(l)>>>(l.ops, l.dims, offset, count, buf);
    RecordLaunchError(hipGetLastError(), "MFMA compression launch");
    return true;
  }
  return false;
}

namespace {

template <typename In, typename Out, int kMode>
bool TransposeTeamFor(const Launch &l, const void *in, void *out, std::uint32_t X, std::uint32_t Y, double *scale) {
  const std::uint64_t threads = static_cast<std::uint64_t>(X) * Y * (l.dims.K / 4);
  if (threads >= (1ull << 32)) return false;
  const unsigned grid = Blocks(threads);
  const auto *i = static_cast<const In *>(in);
  auto *o = static_cast<Out *>(out);
  switch (l.dims.K) {
    case 4: TransposeTeam<In, Out, kMode, 4><<<grid, kBlock, 0, StreamOf(l)>>>(i, o, X, Y, scale); break;
    case 8: TransposeTeam<In, Out, kMode, 8><<<grid, kBlock, 0, StreamOf(l)>>>(i, o, X, Y, scale); break;
    case 16: TransposeTeam<In, Out, kMode, 16><<<grid, kBlock, 0, StreamOf(l)>>>(i, o, X, Y, scale); break;
    case 32: TransposeTeam<In, Out, kMode, 32><<<grid, kBlock, 0, StreamOf(l)>>>(i, o, X, Y, scale); break;
    case 64: TransposeTeam<In, Out, kMode, 64><<<grid, kBlock, 0, StreamOf(l)>>>(i, o, X, Y, scale); break;
    default: return false;
  }
  RecordLaunchError(hipGetLastError(), "vector transpose launch");
  return true;
}

}  // namespace

bool LaunchFastTranspose(const Launch &l, int mode, Real in_real, Real out_real, const void *in, void *out,
                         std::uint32_t X, std::uint32_t Y, double *scale) {
  const std::uint32_t K = l.dims.K;
  if (K % 4 != 0 || K > 64 || (K & (K - 1)) != 0) return false;
  // Below 8 states a vector is under 32 bytes: scattered writes only pay when
  // one side of the transposition is trivial.
  if (K < 8 && X != 1 && Y != 1) return false;
  if (!Aligned16(in) || !Aligned16(out)) return false;
  if (static_cast<std::uint64_t>(X) * Y == 0) return true;
  const bool f_in = in_real == Real::kF32, f_out = out_real == Real::kF32;
  switch (mode) {
    case 0:
      if (f_in != f_out) return false;
      return f_in ? TransposeTeamFor<float, float, 0>(l, in, out, X, Y, scale)
                  : TransposeTeamFor<double, double, 0>(l, in, out, X, Y, scale);
    case 1:
      if (f_in && f_out) return TransposeTeamFor<float, float, 1>(l, in, out, X, Y, scale);
      if (f_in) return TransposeTeamFor<float, double, 1>(l, in, out, X, Y, scale);
      if (f_out) return TransposeTeamFor<double, float, 1>(l, in, out, X, Y, scale);
      return TransposeTeamFor<double, double, 1>(l, in, out, X, Y, scale);
    case 2:
      if (f_in != f_out) return false;
      return f_in ? TransposeTeamFor<float, float, 2>(l, in, out, X, Y, scale)
                  : TransposeTeamFor<double, double, 2>(l, in, out, X, Y, scale);
  }
  return false;
}

}  // namespace thmm
