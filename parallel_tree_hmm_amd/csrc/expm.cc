// Reversible rate matrices: symmetric eigendecomposition and expm(Q t).
#include "expm.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace thmm {
namespace {

// Cyclic Jacobi eigenvalue iteration on the symmetric matrix a (K x K,
// row-major, overwritten): on return a's diagonal holds the eigenvalues and
// u (row-major) the eigenvectors as columns. Each rotation zeroes a[p][q]
// with the numerically stable choice of the rotation angle (the smaller of
// the two roots of t^2 + 2 theta t - 1 = 0).
void JacobiEigen(std::vector<double> &a, std::vector<double> &u, std::size_t K) {
  u.assign(K * K, 0.0);
  for (std::size_t i = 0; i < K; ++i) u[i * K + i] = 1.0;
  double norm = 0.0;
  for (double v : a) norm += v * v;
  if (norm == 0.0) return;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (std::size_t p = 0; p < K; ++p)
      for (std::size_t q = p + 1; q < K; ++q) off += a[p * K + q] * a[p * K + q];
    if (off <= 1e-32 * norm) return;
    for (std::size_t p = 0; p < K; ++p) {
      for (std::size_t q = p + 1; q < K; ++q) {
        const double apq = a[p * K + q];
        if (apq == 0.0) continue;
        const double theta = (a[q * K + q] - a[p * K + p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (std::size_t k = 0; k < K; ++k) {  // columns p, q of a
          const double akp = a[k * K + p], akq = a[k * K + q];
          a[k * K + p] = c * akp - s * akq;
          a[k * K + q] = s * akp + c * akq;
        }
        for (std::size_t k = 0; k < K; ++k) {  // rows p, q of a
          const double apk = a[p * K + k], aqk = a[q * K + k];
          a[p * K + k] = c * apk - s * aqk;
          a[q * K + k] = s * apk + c * aqk;
        }
        for (std::size_t k = 0; k < K; ++k) {  // eigenvector columns
          const double ukp = u[k * K + p], ukq = u[k * K + q];
          u[k * K + p] = c * ukp - s * ukq;
          u[k * K + q] = s * ukp + c * ukq;
        }
      }
    }
  }
  throw std::invalid_argument("rate-matrix eigendecomposition did not converge");
}

}  // namespace

ReversibleRates DecomposeReversible(const double *rate, const double *stationary, std::size_t K) {
  if (rate == nullptr || stationary == nullptr) throw std::invalid_argument("rate matrix or stationary law is null");
  if (K == 0) throw std::invalid_argument("rate matrix needs at least one state");
  double flux = 0.0;
  for (std::size_t i = 0; i < K; ++i) {
    if (!(std::isfinite(stationary[i]) && stationary[i] > 0.0))
      throw std::invalid_argument("stationary probabilities must be positive and finite");
    for (std::size_t j = 0; j < K; ++j) {
      const double q = rate[i * K + j];
      if (!std::isfinite(q)) throw std::invalid_argument("rate matrix entries must be finite");
      if (i != j && q < 0.0) throw std::invalid_argument("off-diagonal rates must be nonnegative");
      if (i != j) flux = std::max(flux, stationary[i] * q);
    }
  }
  for (std::size_t i = 0; i < K; ++i)
    for (std::size_t j = i + 1; j < K; ++j)
      if (std::fabs(stationary[i] * rate[i * K + j] - stationary[j] * rate[j * K + i]) > 1e-10 * flux)
        throw std::invalid_argument("rate matrix is not reversible with respect to the stationary law");
  std::vector<double> root(K);
  for (std::size_t i = 0; i < K; ++i) root[i] = std::sqrt(stationary[i]);
  std::vector<double> s(K * K);
  for (std::size_t i = 0; i < K; ++i)
    for (std::size_t j = 0; j < K; ++j) {
      const double sij = root[i] * rate[i * K + j] / root[j];
      const double sji = root[j] * rate[j * K + i] / root[i];
      s[i * K + j] = 0.5 * (sij + sji);
    }
  std::vector<double> u;
  JacobiEigen(s, u, K);
  ReversibleRates out;
  out.states = K;
  out.lambda.resize(K);
  out.V.resize(K * K);
  out.V_inv.resize(K * K);
  for (std::size_t k = 0; k < K; ++k) out.lambda[k] = s[k * K + k];
  for (std::size_t i = 0; i < K; ++i)
    for (std::size_t k = 0; k < K; ++k) {
      out.V[i * K + k] = u[i * K + k] / root[i];
      out.V_inv[k * K + i] = u[i * K + k] * root[i];
    }
  return out;
}

void CheckBranchLengths(const double *branch_lengths, std::size_t edges) {
  if (edges != 0 && branch_lengths == nullptr) throw std::invalid_argument("branch lengths are null");
  for (std::size_t e = 0; e < edges; ++e)
    if (!(std::isfinite(branch_lengths[e]) && branch_lengths[e] >= 0.0))
      throw std::invalid_argument("branch lengths must be finite and nonnegative");
}

void ExpmReversible(const ReversibleRates &r, const double *branch_lengths, std::size_t edges, double *out) {
  const std::size_t K = r.states;
  std::vector<double> w(K);
  for (std::size_t e = 0; e < edges; ++e) {
    for (std::size_t k = 0; k < K; ++k) w[k] = std::exp(r.lambda[k] * branch_lengths[e]);
    double *p = out + e * K * K;
    for (std::size_t i = 0; i < K; ++i)
      for (std::size_t j = 0; j < K; ++j) {
        double sum = 0.0;
        for (std::size_t k = 0; k < K; ++k) sum += r.V[i * K + k] * w[k] * r.V_inv[k * K + j];
        p[i * K + j] = std::max(sum, 0.0);
      }
  }
}

}  // namespace thmm
