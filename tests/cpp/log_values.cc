// ctypes entry for the tests: LogOfFloat (parallel_tree_hmm_amd/csrc/log_f32.h)
// on the host, with the MAP kernels' rule for non-positive factors (-inf).
// Built by tests/test_map_log.py with g++ -O2 -ffp-contract=off.
#include <cmath>
#include <cstddef>

#include "log_f32.h"

extern "C" void log_of_floats(const float *x, double *out, std::size_t n) {
  for (std::size_t i = 0; i < n; ++i) out[i] = x[i] > 0.0f ? thmm::LogOfFloat(x[i], thmm::kLogF32Table) : -INFINITY;
}
