// FP64 natural logarithm of a positive finite FP32 value, for the max-product
// (MAP) kernels, whose log-domain inputs are the logs of FP32 potentials.
//
// Method (table reduction; Tang, "Table-driven implementation of the
// logarithm function in IEEE floating-point arithmetic", ACM TOMS 16(4),
// 1990): x = 2^e z with z in [1, 2) and 24 significant bits; with
// i = the top 7 fraction bits of z and invc_i ~ 1 / c_i from a 128-entry
// table (24-bit floats),
//   r = z invc_i - 1                      (exact: a 48-bit product minus 1)
//   log x = k ln 2 - log(invc_i) + log1p(r),   |r| <= 2^-7,
// where k = e, or e + 1 for i = 127 (z near 2 is taken as z / 2 near 1,
// invc = 1/2, so arguments just below a power of two keep full relative
// accuracy). log1p(r) = r + r^2 P(r) with the Taylor terms to r^9
// (truncation below 2^-66 relative). The head k ln2_hi - log(invc)_hi + r is
// summed with two error-free TwoSums; the tails (k ln2_lo, the table's
// correction, both TwoSum errors, r^2 P(r)) are added before the final
// rounding. Every operation is an IEEE FP64 add, multiply or fused
// multiply-add written out explicitly (contraction is off), so host and
// device evaluations give the same bits. Errors: tests/test_map_log.py.
//
// Usable from device code (hipcc) and host C++ (the host check).
#ifndef TREE_HMM_AMD_LOG_F32_H_
#define TREE_HMM_AMD_LOG_F32_H_

#include <cmath>
#include <cstdint>
#include <cstring>

#include "log_f32_table.h"

#if defined(__HIPCC__)
#define THMM_HOST_DEVICE __host__ __device__
#else
#define THMM_HOST_DEVICE
#endif

namespace thmm {

struct LogF32Entry {
  double hi;   // -log(invc), rounded to FP64
  float lo;    // -log(invc) - hi
  float invc;  // 1 / c_i with 24 significant bits
};

// The table for host code; device code reads a copy (kernels stage it in
// LDS, see kernels.hip).
static const LogF32Entry kLogF32Table[128] = THMM_LOG_F32_TABLE;

namespace logf32_detail {

THMM_HOST_DEVICE inline std::uint64_t Bits(double x) {
  std::uint64_t u;
  std::memcpy(&u, &x, sizeof u);
  return u;
}
THMM_HOST_DEVICE inline double FromBits(std::uint64_t u) {
  double x;
  std::memcpy(&x, &u, sizeof x);
  return x;
}

}  // namespace logf32_detail

// log(x) in FP64 for a positive finite float x (+inf for +inf), with the
// reduction table `table` (a copy of kLogF32Table).
THMM_HOST_DEVICE inline double LogOfFloat(float xf, const LogF32Entry *table) {
#pragma clang fp contract(off)
  using namespace logf32_detail;
  const double x = static_cast<double>(xf);
  if (x == static_cast<double>(INFINITY)) return x;
  const std::uint64_t bits = Bits(x);  // a normal double: FP32 subnormals are normal here
  const int e = static_cast<int>((bits >> 52) & 0x7FF) - 1023;
  const int i = static_cast<int>((bits >> 45) & 127);
  const double z = FromBits((bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  const LogF32Entry t = table[i];
  const double k = static_cast<double>(e + (i == 127 ? 1 : 0));
  const double r = std::fma(z, static_cast<double>(t.invc), -1.0);
  // Head: k ln2_hi (exact) + hi, then + r, each with its rounding error.
  const double a = k * THMM_LN2_HI;
  const double s = a + t.hi;
  const double sb = s - a;
  const double s_err = (a - (s - sb)) + (t.hi - sb);
  const double u = s + r;
  const double ub = u - s;
  const double u_err = (s - (u - ub)) + (r - ub);
  // log1p(r) - r = r^2 (-1/2 + r (1/3 + r (-1/4 + ... + r / 9))).
  double p = 1.0 / 9.0;
  p = std::fma(p, r, -1.0 / 8.0);
  p = std::fma(p, r, 1.0 / 7.0);
  p = std::fma(p, r, -1.0 / 6.0);
  p = std::fma(p, r, 1.0 / 5.0);
  p = std::fma(p, r, -1.0 / 4.0);
  p = std::fma(p, r, 1.0 / 3.0);
  p = std::fma(p, r, -0.5);
  const double r2 = r * r;
  const double tail = std::fma(k, THMM_LN2_LO, static_cast<double>(t.lo)) + (s_err + u_err) + r2 * p;
  return u + tail;
}

}  // namespace thmm

#endif  // TREE_HMM_AMD_LOG_F32_H_
