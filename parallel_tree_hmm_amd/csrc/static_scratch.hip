// Workspace-free log partition: see static_scratch.h.
//
// The scratch array is one file-scope __device__ array; every GPU the
// library runs on holds its own copy until the process ends. Host code
// never passes the array's name to a kernel (in host code that name is a
// host-side stand-in): the first call on a device, with that device
// current, looks up the device address with hipGetSymbolAddress and keeps it
// per device, and every launch, copy and pointer below uses that address.
#include <hip/hip_runtime.h>

#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>

#include "static_scratch.h"

namespace thmm {
namespace {

constexpr std::size_t kScratchBytes = THMM_STATIC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch[kMaxDevices] = {};

void Hip(hipError_t status, const char *what) {
  if (status != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
}

// The device address of `scratch` on `device` (which is current).
unsigned char *ScratchOf(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch[device] == nullptr) {
    void *ptr = nullptr;
    Hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "GPU runtime GetSymbolAddress");
    if (ptr == nullptr) throw std::runtime_error("GPU runtime GetSymbolAddress returned a null address");
    g_scratch[device] = static_cast<unsigned char *>(ptr);
  }
  return g_scratch[device];
}

std::size_t Mul(std::size_t a, std::size_t b) {
  if (a != 0 && b > std::numeric_limits<std::size_t>::max() / a)
    throw std::length_error("static scratch size overflows size_t");
  return a * b;
}

std::size_t Aligned(std::size_t bytes) {
  if (bytes > std::numeric_limits<std::size_t>::max() - 255) throw std::length_error("static scratch size overflows");
  return (bytes + 255) & ~static_cast<std::size_t>(255);
}

// Byte offsets of the pieces, each 256-byte aligned.
struct Layout {
  std::size_t rakes, combines, absorbs, compresses, roots, nodes, node_scale, paths, path_scale, branches,
      branch_scale, total;
};

Layout LayoutOf(const Plan &plan, std::size_t K, std::size_t B, std::size_t s) {
  Layout l{};
  std::size_t at = 0;
  const auto take = [&](std::size_t bytes) {
    const std::size_t here = at;
    at = Aligned(here + bytes);
    if (at < here) throw std::length_error("static scratch size overflows");
    return here;
  };
  const std::size_t N = plan.num_nodes, R = plan.num_branches, C = plan.num_compressions();
  l.rakes = take(Mul(plan.rakes.size(), sizeof(DeviceRake)));
  l.combines = take(Mul(plan.combinations.size(), sizeof(DeviceCombine)));
  l.absorbs = take(Mul(plan.absorptions.size(), sizeof(DeviceAbsorb)));
  l.compresses = take(Mul(C, sizeof(DeviceCompress)));
  l.roots = take(Mul(plan.roots.size(), sizeof(std::uint32_t)));
  l.nodes = take(Mul(Mul(N, B), Mul(K, s)));
  l.node_scale = take(Mul(Mul(N, B), sizeof(double)));
  l.paths = take(Mul(Mul(C, B), Mul(Mul(K, K), s)));
  l.path_scale = take(Mul(Mul(C, B), Mul(K, sizeof(double))));
  l.branches = take(Mul(Mul(R, B), Mul(K, s)));
  l.branch_scale = take(Mul(Mul(R, B), sizeof(double)));
  l.total = at;
  return l;
}

template <typename V>
void Copy(unsigned char *base, std::size_t offset, const std::vector<V> &src, hipStream_t stream) {
  if (src.empty()) return;
  Hip(hipMemcpyAsync(base + offset, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice, stream),
      "GPU runtime MemcpyAsync static topology upload");
}

}  // namespace

std::size_t StaticScratchBytes(const Plan &plan, std::size_t states, std::size_t batch, Real real) {
  return LayoutOf(plan, states, batch, real == Real::kF32 ? sizeof(float) : sizeof(double)).total;
}

void LogPartitionStatic(const Plan &plan, const RenamedOps &ops, std::size_t states, std::size_t batch, Real real,
                        const void *nodes, const void *edges, double *log_z, void *stream, int device) {
  if (states == 0 || batch == 0) throw std::invalid_argument("accelerator state and batch counts must be nonzero");
  if (states > 64) throw std::invalid_argument("state count exceeds the accelerator limit of 64");
  if (real != Real::kF32 && real != Real::kF64) throw std::invalid_argument("unknown precision");
  if (nodes == nullptr || (plan.num_edges != 0 && edges == nullptr) || log_z == nullptr)
    throw std::invalid_argument("device input or output is null");
  if (Mul(plan.num_nodes, batch) > std::numeric_limits<std::uint32_t>::max())
    throw std::length_error("node-batch count exceeds uint32_t");
  const std::size_t s = real == Real::kF32 ? sizeof(float) : sizeof(double);
  const Layout layout = LayoutOf(plan, states, batch, s);
  if (layout.total > kScratchBytes)
    throw std::length_error("thmm_log_partition_device_static needs " + std::to_string(layout.total) +
                            " bytes of contraction state, above its static maximum of " +
                            std::to_string(kScratchBytes) + " bytes (use a reserved workspace)");
  Hip(hipSetDevice(device), "GPU runtime SetDevice");
  unsigned char *base = ScratchOf(device);
  const hipStream_t s_ = (stream == nullptr || stream == THMM_HIP_NULL_STREAM) ? nullptr
                                                                               : static_cast<hipStream_t>(stream);
  Copy(base, layout.rakes, ops.rakes, s_);
  Copy(base, layout.combines, ops.combines, s_);
  Copy(base, layout.absorbs, ops.absorbs, s_);
  Copy(base, layout.compresses, ops.compresses, s_);
  Copy(base, layout.roots, plan.roots, s_);

  Launch l{};
  l.real = real;
  l.dims = {static_cast<std::uint32_t>(states), static_cast<std::uint32_t>(batch),
            static_cast<std::uint32_t>(plan.num_edges), static_cast<std::uint32_t>(plan.num_nodes),
            static_cast<std::uint32_t>(plan.roots.size())};
  l.ops = {reinterpret_cast<const DeviceRake *>(base + layout.rakes),
           reinterpret_cast<const DeviceCombine *>(base + layout.combines),
           reinterpret_cast<const DeviceAbsorb *>(base + layout.absorbs),
           reinterpret_cast<const DeviceCompress *>(base + layout.compresses),
           reinterpret_cast<const std::uint32_t *>(base + layout.roots), nullptr};
  l.stream = s_;
  l.tuning = Tuning{};
  LinearBuffers buf{};
  buf.nodes = base + layout.nodes;
  buf.node_scale = reinterpret_cast<double *>(base + layout.node_scale);
  buf.edges = edges;
  buf.paths = base + layout.paths;
  buf.path_scale = reinterpret_cast<double *>(base + layout.path_scale);
  buf.branches = base + layout.branches;
  buf.branch_scale = reinterpret_cast<double *>(base + layout.branch_scale);
  // Every value the pass reads is written earlier in the same pass (nodes by
  // the input transform, branches by rakes, paths by compressions), so
  // nothing needs clearing between calls.
  LaunchInitNodes(l, real, nodes, buf.nodes, buf.node_scale);
  for (const thmm_primitive_batch &b : plan.batches) LaunchForwardLevel(l, b.primitive, b.offset, b.count, buf);
  LaunchFinish(l, buf, log_z);
  const char *error = LastLaunchError();
  if (error != nullptr && error[0] != '\0') throw std::runtime_error(error);
}

}  // namespace thmm
