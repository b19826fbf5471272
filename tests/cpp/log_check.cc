// Host check of LogOfFloat (parallel_tree_hmm_amd/csrc/log_f32.h), the FP64
// log of FP32 values used by the MAP kernels, over positive finite floats
// with bit patterns first, first + stride, ... (stride 1: all of them):
//   * agreement, bit for bit, with the C library's log of the same double,
//     the function the FP64 oracle (oracle/oracle.cc) takes logs with;
//   * the error against the x87 extended-precision logl, in units in the
//     last place of the double result.
// Built by tests/test_map_log.py with g++ -O2 -ffp-contract=off, the same
// FP64 operation sequence the device code compiles to.
// Usage: log_check STRIDE FIRST [THREADS]
// Prints: values, differing from log, largest |error| in ulps, results
// further than 0.5 ulp (+ 2^-10) from logl, log's results likewise.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "log_f32.h"

namespace {

struct Tally {
  std::uint64_t values = 0, differ = 0, ours_far = 0, libm_far = 0;
  double worst = 0.0;
};

double UlpError(double y, long double ref) {
  const double ulp = std::nextafter(std::fabs(y), INFINITY) - std::fabs(y);
  return static_cast<double>((static_cast<long double>(y) - ref) / ulp);
}

void Sweep(std::uint64_t first, std::uint64_t stride, std::uint64_t lane, std::uint64_t lanes, Tally *out) {
  const std::uint64_t last = 0x7F7FFFFFull;  // largest finite float
  Tally t;
  for (std::uint64_t b = first + lane * stride; b <= last; b += lanes * stride) {
    if (b == 0) continue;
    const std::uint32_t bits = static_cast<std::uint32_t>(b);
    float x;
    std::memcpy(&x, &bits, sizeof x);
    const double ours = thmm::LogOfFloat(x, thmm::kLogF32Table);
    const double libm = std::log(static_cast<double>(x));
    const long double ref = logl(static_cast<long double>(x));
    ++t.values;
    if (std::memcmp(&ours, &libm, sizeof ours) != 0) ++t.differ;
    if (ours == 0.0 && libm == 0.0) continue;
    const double e = UlpError(ours, ref), l = UlpError(libm, ref);
    if (std::fabs(e) > t.worst) t.worst = std::fabs(e);
    if (std::fabs(e) > 0.5 + 1.0 / 1024) ++t.ours_far;
    if (std::fabs(l) > 0.5 + 1.0 / 1024) ++t.libm_far;
  }
  *out = t;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: log_check STRIDE FIRST [THREADS]\n");
    return 2;
  }
  const std::uint64_t stride = std::strtoull(argv[1], nullptr, 10);
  const std::uint64_t first = std::strtoull(argv[2], nullptr, 10);
  const unsigned threads = argc > 3 ? static_cast<unsigned>(std::atoi(argv[3])) : 1;
  if (stride == 0 || threads == 0) return 2;
  std::vector<Tally> tallies(threads);
  std::vector<std::thread> pool;
  for (unsigned i = 0; i < threads; ++i) pool.emplace_back(Sweep, first, stride, i, threads, &tallies[i]);
  for (std::thread &th : pool) th.join();
  Tally sum;
  for (const Tally &t : tallies) {
    sum.values += t.values;
    sum.differ += t.differ;
    sum.ours_far += t.ours_far;
    sum.libm_far += t.libm_far;
    if (t.worst > sum.worst) sum.worst = t.worst;
  }
  std::printf("%llu %llu %.6f %llu %llu\n", static_cast<unsigned long long>(sum.values),
              static_cast<unsigned long long>(sum.differ), sum.worst, static_cast<unsigned long long>(sum.ours_far),
              static_cast<unsigned long long>(sum.libm_far));
  return 0;
}
