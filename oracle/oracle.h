/*
 * FP64 CPU oracle for the batched tree-HMM path. TEST INFRASTRUCTURE ONLY:
 * it is linked by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg as the checker, and never by the product library.
 *
 * It restates the reference CPU inference of joaospinto/parallel_tree_hmm
 * (src/inference.cc) over a rake-compress plan given as flat arrays
 * (thmm_plan_arrays from include/tree_hmm_amd.h: only the struct layout is
 * shared). Every function names the reference lines it follows.
 *
 * Forests: where the reference has one root (plan.root()), these functions
 * treat every root the same way, returning one log partition per root.
 *
 * Return codes: 0 ok, 1 invalid argument, 4 domain error (matching
 * thmm_status).
 */
#ifndef TREE_HMM_AMD_ORACLE_H_
#define TREE_HMM_AMD_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#include "../include/tree_hmm_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Unscaled sum-product Z per root (src/inference.cc:148-245,885-890). */
int oracle_partition(const thmm_plan_arrays *plan, size_t states,
                     const double *nodes, const double *edges, double *z);

/* Scaled sum-product log Z per root (src/inference.cc:697-836,892-903). */
int oracle_log_partition(const thmm_plan_arrays *plan, size_t states,
                         const double *nodes, const double *edges,
                         double *log_z);

/* Log-domain forward + analytic reverse: log Z per root, node marginals
 * [node, state], edge marginals [edge, parent, child]
 * (src/inference.cc:247-562,905-917). */
int oracle_marginals(const thmm_plan_arrays *plan, size_t states,
                     const double *nodes, const double *edges, double *log_z,
                     double *node_marginals, double *edge_marginals);

/* Max-product MAP: log weight per root and states [node]
 * (src/inference.cc:564-695,919-927). */
int oracle_map(const thmm_plan_arrays *plan, size_t states,
               const double *nodes, const double *edges, double *log_weight,
               uint32_t *assignment);

/* Posterior sample from one variate per node
 * (src/inference.cc:128-146,345-412,929-939). boundary_distance (may be
 * NULL) receives, per node, the distance from its variate to the nearest
 * interior boundary of the inverse CDF it was drawn from. */
int oracle_sample(const thmm_plan_arrays *plan, size_t states,
                  const double *nodes, const double *edges,
                  const double *uniforms, uint32_t *assignment,
                  double *boundary_distance);

/* Exhaustive enumeration over states^nodes assignments (tiny inputs only):
 * Z, node and edge marginals (tests/inference_test.cc:55-97) and the
 * maximum-weight assignment, first in enumeration order with node 0 varying
 * fastest (tests/inference_test.cc:119-144). Any output may be NULL. */
int oracle_brute_force(const thmm_plan_arrays *plan, size_t states,
                       const double *nodes, const double *edges, double *z,
                       double *node_marginals, double *edge_marginals,
                       double *max_weight, uint32_t *max_assignment);

/* Product of potentials along one assignment
 * (tests/inference_test.cc:99-117). */
double oracle_assignment_weight(const thmm_plan_arrays *plan, size_t states,
                                const double *nodes, const double *edges,
                                const uint32_t *assignment);

/* Exhaustive conditional sampler: visits nodes in the plan's expansion
 * order (roots, then rake leaves / compression middles of the batches in
 * reverse) and draws each from its exact conditional given the nodes drawn
 * before it, by enumeration, with the inverse-CDF rule of
 * src/inference.cc:128-146. Tiny inputs only. */
int oracle_brute_conditional_sample(const thmm_plan_arrays *plan,
                                    size_t states, const double *nodes,
                                    const double *edges,
                                    const double *uniforms,
                                    uint32_t *assignment);

/* Runs oracle_marginals over `items` independent batch items on `threads`
 * host threads (static partition): nodes [items, node, state], shared edges,
 * outputs log_z [items, roots], node marginals [items, node, state], edge
 * marginals [items, edge, parent, child]. CPU-baseline helper. */
int oracle_marginals_batched(const thmm_plan_arrays *plan, size_t states,
                             size_t items, const double *nodes,
                             const double *edges, double *log_z,
                             double *node_marginals, double *edge_marginals,
                             size_t threads);

#ifdef __cplusplus
}
#endif

#endif /* TREE_HMM_AMD_ORACLE_H_ */
