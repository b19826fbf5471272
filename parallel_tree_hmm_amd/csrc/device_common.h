// Device helpers shared by the kernel translation units.
#ifndef TREE_HMM_AMD_DEVICE_COMMON_H_
#define TREE_HMM_AMD_DEVICE_COMMON_H_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "launch.h"

namespace thmm {

// Keeps the first failed launch of the calling host thread for
// LastLaunchError().
void RecordLaunchError(hipError_t status, const char *what);

// Launches a level on the coalesced/MFMA kernels when they cover (K,
// element type, primitive); returns false to leave it to the generic ones.
bool LaunchFastLevel(const Launch &l, bool reverse, std::uint32_t primitive, std::uint32_t offset,
                     std::uint32_t count, const LinearBuffers &buf);

// Transposes K-vectors [X][Y][K] -> [Y][X][K] on the coalesced team kernel
// (mode 0 copy, 1 linear -> shifted logs, 2 logs -> shifted logs, 3 logs ->
// max-shifted exponentials, the latter three writing FP64 scales); returns
// false when (K, types, shape) are left to the generic kernels. With `only`
// (mode 1) just the vectors y listed there are moved, for every x.
bool LaunchFastTranspose(const Launch &l, int mode, Real in_real, Real out_real, const void *in, void *out,
                         std::uint32_t X, std::uint32_t Y, double *scale, const NodeList *only = nullptr);

namespace dev {

constexpr double kNegInf = -__builtin_huge_val();

__device__ inline float Max(float a, float b) { return fmaxf(a, b); }
__device__ inline double Max(double a, double b) { return fmax(a, b); }
__device__ inline float Exp(float x) { return expf(x); }
__device__ inline double Exp(double x) { return exp(x); }
// exp(y) for y <= 0 (or -inf): the hardware base-2 exponential of y log2(e),
// relative error about |y| 6e-8 from the rounded product; used where log
// potentials are shifted by their row maximum.
__device__ inline float ExpNonPositive(float y) { return __builtin_amdgcn_exp2f(y * 1.44269504088896341f); }
__device__ inline double ExpNonPositive(double y) { return exp(y); }

template <typename T> __device__ inline T NegInf() { return static_cast<T>(kNegInf); }

template <int W, typename V> __device__ inline V GroupMax(V v) {
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) v = Max(v, __shfl_xor(v, o, W));
  return v;
}

// Storage of slot `slot` for batch item b: edges are batch-shared, composed
// paths per item.
template <typename T>
__device__ inline const T *Slot(const LinearBuffers &buf, const Dims &d, std::uint32_t slot, std::uint32_t b) {
  const std::uint64_t kk = static_cast<std::uint64_t>(d.K) * d.K;
  return slot < d.E ? static_cast<const T *>(buf.edges) + slot * kk
                    : static_cast<const T *>(buf.paths) + (static_cast<std::uint64_t>(slot - d.E) * d.B + b) * kk;
}

// FP64 log scale of row `row` of a slot.
__device__ inline double RowScale(const LinearBuffers &buf, const Dims &d, std::uint32_t slot, std::uint32_t b,
                                  std::uint32_t row) {
  if (slot < d.E) return buf.edge_scale ? buf.edge_scale[static_cast<std::uint64_t>(slot) * d.K + row] : 0.0;
  return buf.path_scale[(static_cast<std::uint64_t>(slot - d.E) * d.B + b) * d.K + row];
}

// Adjoint storage of a slot: edge marginals are written batch-major
// [B][E][K][K] straight into the output; composed slots are internal.
template <typename T>
__device__ inline T *SlotAdjoint(const LinearBuffers &buf, const Dims &d, std::uint32_t slot, std::uint32_t b) {
  const std::uint64_t kk = static_cast<std::uint64_t>(d.K) * d.K;
  return slot < d.E ? static_cast<T *>(buf.edge_out) + (static_cast<std::uint64_t>(b) * d.E + slot) * kk
                    : static_cast<T *>(buf.path_adj) + (static_cast<std::uint64_t>(slot - d.E) * d.B + b) * kk;
}

// Adjoint of a rake's branch message: the branch's own adjoint, or, when the
// executor keeps no branch adjoints (every branch is absorbed into the rake's
// parent and combination/absorption adjoints are copies), the parent's node
// adjoint.
template <typename T>
__device__ inline const T *RakeAdjoint(const LinearBuffers &buf, const Dims &d, const DeviceRake &op, std::uint32_t b) {
  if (buf.branch_adj != nullptr)
    return static_cast<const T *>(buf.branch_adj) + (static_cast<std::uint64_t>(op.branch) * d.B + b) * d.K;
  return static_cast<const T *>(buf.node_adj) + (static_cast<std::uint64_t>(op.parent) * d.B + b) * d.K;
}

// Four consecutive elements, moved with 16-byte (float) or 2 x 16-byte
// (double) vector memory instructions.
template <typename T> struct Quad {
  T v[4];
};

template <typename T> __device__ inline Quad<T> LoadQuad(const T *p) {
  Quad<T> q;
  if constexpr (sizeof(T) == 4) {
    const float4 x = *reinterpret_cast<const float4 *>(p);
    q.v[0] = x.x, q.v[1] = x.y, q.v[2] = x.z, q.v[3] = x.w;
  } else {
    const double2 x = *reinterpret_cast<const double2 *>(p);
    const double2 y = *reinterpret_cast<const double2 *>(p + 2);
    q.v[0] = x.x, q.v[1] = x.y, q.v[2] = y.x, q.v[3] = y.y;
  }
  return q;
}

template <typename T> __device__ inline void StoreQuad(T *p, const Quad<T> &q) {
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4 *>(p) = make_float4(q.v[0], q.v[1], q.v[2], q.v[3]);
  } else {
    *reinterpret_cast<double2 *>(p) = make_double2(q.v[0], q.v[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(q.v[2], q.v[3]);
  }
}

// Streaming forms for data read or written exactly once by a level (edge
// slots, edge marginals): non-temporal, so they do not displace reusable
// lines in the caches.
typedef float StreamF4 __attribute__((ext_vector_type(4)));
typedef double StreamD2 __attribute__((ext_vector_type(2)));

template <typename T> __device__ inline Quad<T> LoadQuadStream(const T *p) {
  Quad<T> q;
  if constexpr (sizeof(T) == 4) {
    const StreamF4 x = __builtin_nontemporal_load(reinterpret_cast<const StreamF4 *>(p));
    q.v[0] = x[0], q.v[1] = x[1], q.v[2] = x[2], q.v[3] = x[3];
  } else {
    const StreamD2 x = __builtin_nontemporal_load(reinterpret_cast<const StreamD2 *>(p));
    const StreamD2 y = __builtin_nontemporal_load(reinterpret_cast<const StreamD2 *>(p + 2));
    q.v[0] = x[0], q.v[1] = x[1], q.v[2] = y[0], q.v[3] = y[1];
  }
  return q;
}

template <typename T> __device__ inline void StoreQuadStream(T *p, const Quad<T> &q) {
  if constexpr (sizeof(T) == 4) {
    const StreamF4 x = {q.v[0], q.v[1], q.v[2], q.v[3]};
    __builtin_nontemporal_store(x, reinterpret_cast<StreamF4 *>(p));
  } else {
    const StreamD2 x = {q.v[0], q.v[1]}, y = {q.v[2], q.v[3]};
    __builtin_nontemporal_store(x, reinterpret_cast<StreamD2 *>(p));
    __builtin_nontemporal_store(y, reinterpret_cast<StreamD2 *>(p + 2));
  }
}

// A dispatch's grid size is a 32-bit count of work-items (blocks times
// their threads). Kernels over arrays that can exceed it run a capped grid
// (GridStrideBlocks) and step through the array by the grid's size.
__device__ inline std::uint64_t GridThread() {
  return static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}
__device__ inline std::uint64_t GridStride() { return static_cast<std::uint64_t>(gridDim.x) * blockDim.x; }

// Four consecutive values of a slot from element `at` (a multiple of 4) and
// one value at `at`: edge slots widened from FP32 when buf.edges_f32 is set.
template <typename T>
__device__ inline Quad<T> LoadSlotQuad(const LinearBuffers &buf, const Dims &d, std::uint32_t slot, std::uint32_t b,
                                       std::uint64_t at) {
  if constexpr (std::is_same<T, double>::value) {
    if (slot < d.E && buf.edges_f32) {
      const Quad<float> f =
          LoadQuad(static_cast<const float *>(buf.edges) + static_cast<std::uint64_t>(slot) * d.K * d.K + at);
      Quad<double> q;
#pragma unroll
      for (int r = 0; r < 4; ++r) q.v[r] = static_cast<double>(f.v[r]);
      return q;
    }
  }
  return LoadQuad(Slot<T>(buf, d, slot, b) + at);
}

template <typename T>
__device__ inline T SlotValue(const LinearBuffers &buf, const Dims &d, std::uint32_t slot, std::uint32_t b,
                              std::uint64_t at) {
  if constexpr (std::is_same<T, double>::value) {
    if (slot < d.E && buf.edges_f32)
      return static_cast<double>(
          static_cast<const float *>(buf.edges)[static_cast<std::uint64_t>(slot) * d.K * d.K + at]);
  }
  return Slot<T>(buf, d, slot, b)[at];
}

// Where a slot's adjoint goes: composed-path adjoints in the element type,
// edge marginals in the element type or, with buf.edge_out_half, as IEEE
// binary16 (rounded to nearest even; magnitudes below 2^-14 are stored as
// subnormals, with an absolute resolution of 2^-24).
typedef _Float16 Half4 __attribute__((ext_vector_type(4)));

template <typename T> struct AdjointTarget {
  T *wide;
  _Float16 *half;
};

template <typename T>
__device__ inline AdjointTarget<T> SlotAdjointTarget(const LinearBuffers &buf, const Dims &d, std::uint32_t slot,
                                                     std::uint32_t b) {
  if (slot < d.E && buf.edge_out_half) {
    const std::uint64_t kk = static_cast<std::uint64_t>(d.K) * d.K;
    return {nullptr, static_cast<_Float16 *>(buf.edge_out) + (static_cast<std::uint64_t>(b) * d.E + slot) * kk};
  }
  return {SlotAdjoint<T>(buf, d, slot, b), nullptr};
}

// Four consecutive adjoint values at element offset `at` (a multiple of 4).
template <typename T, bool kStream = true>
__device__ inline void StoreAdjointQuad(const AdjointTarget<T> &t, std::uint64_t at, const Quad<T> &q) {
  if (t.half != nullptr) {
    const Half4 h = {static_cast<_Float16>(q.v[0]), static_cast<_Float16>(q.v[1]), static_cast<_Float16>(q.v[2]),
                     static_cast<_Float16>(q.v[3])};
    if (kStream)
      __builtin_nontemporal_store(h, reinterpret_cast<Half4 *>(t.half + at));
    else
      *reinterpret_cast<Half4 *>(t.half + at) = h;
  } else if (kStream) {
    StoreQuadStream(t.wide + at, q);
  } else {
    StoreQuad(t.wide + at, q);
  }
}

template <typename T> __device__ inline void StoreAdjointValue(const AdjointTarget<T> &t, std::uint64_t at, T v) {
  if (t.half != nullptr)
    t.half[at] = static_cast<_Float16>(v);
  else
    t.wide[at] = v;
}

}  // namespace dev
}  // namespace thmm

#endif  // TREE_HMM_AMD_DEVICE_COMMON_H_
