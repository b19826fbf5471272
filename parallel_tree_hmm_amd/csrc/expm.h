// Transition matrices P(t) = expm(Q t) of a reversible rate matrix (host C++).
//
// The reference takes edge factors as caller-staged values
// (include/tree_hmm/accelerator.h:86-96, CategoricalInputUpdate::kFactors);
// a phylogenetic caller forms them as expm(Q t_e) for every branch length.
// For Q reversible with respect to pi (pi_i Q_ij = pi_j Q_ji),
// S = D^1/2 Q D^-1/2 (D = diag(pi)) is symmetric, S = U diag(lambda) U^T,
// and Q = V diag(lambda) V^-1 with V = D^-1/2 U and V^-1 = U^T D^1/2, so
//   P(t)[i][j] = sum_k V[i][k] exp(lambda_k t) V^-1[k][j].
#ifndef TREE_HMM_AMD_EXPM_H_
#define TREE_HMM_AMD_EXPM_H_

#include <cstddef>
#include <vector>

namespace thmm {

struct ReversibleRates {
  std::size_t states = 0;
  std::vector<double> lambda;   // [K]
  std::vector<double> V;        // [K][K], row-major
  std::vector<double> V_inv;    // [K][K], row-major
};

// Eigendecomposition of Q (row-major K x K) through its symmetric form
// (cyclic Jacobi rotations, FP64). Throws invalid_argument when pi is not
// positive and finite, Q is not finite, has a negative off-diagonal rate,
// or violates detailed balance beyond 1e-10 of the largest flux pi_i Q_ij.
ReversibleRates DecomposeReversible(const double *rate, const double *stationary, std::size_t states);

// Throws invalid_argument unless every branch length is finite and >= 0.
void CheckBranchLengths(const double *branch_lengths, std::size_t edges);

// out[e][i][j] = P(t_e)[i][j], FP64; entries below zero (rounding of
// entries that are zero in exact arithmetic) are set to zero.
void ExpmReversible(const ReversibleRates &rates, const double *branch_lengths, std::size_t edges, double *out);

}  // namespace thmm

#endif  // TREE_HMM_AMD_EXPM_H_
