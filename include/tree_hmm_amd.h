/*
 * tree_hmm_amd.h: the C-ABI of the MI355X-native batched tree-HMM inference
 * library (libtree_hmm_amd.so).
 *
 * Every entry point takes plain pointers and sizes and returns a
 * thmm_status. The status codes map one to one onto the exception classes
 * the reference C++ API throws (invalid_argument, logic_error, length_error,
 * domain_error, runtime_error); thmm_last_error() returns the message of the
 * most recent failure on the calling thread. include/tree_hmm_amd/rocm.h
 * rebuilds the reference's tree_hmm::rocm C++ surface on top of this file.
 *
 * Citations name files of joaospinto/parallel_tree_hmm.
 */
#ifndef TREE_HMM_AMD_H_
#define TREE_HMM_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status                                                                   */
/* ------------------------------------------------------------------------ */

typedef enum thmm_status {
  THMM_OK = 0,
  THMM_INVALID_ARGUMENT = 1, /* std::invalid_argument */
  THMM_LOGIC_ERROR = 2,      /* std::logic_error */
  THMM_LENGTH_ERROR = 3,     /* std::length_error */
  THMM_DOMAIN_ERROR = 4,     /* std::domain_error */
  THMM_RUNTIME_ERROR = 5     /* std::runtime_error (HIP failures) */
} thmm_status;

/* Message of the last failing call on this thread ("" if none). Replaces the
 * what() strings of the reference's exceptions
 * (src/gpu_backend_impl.inc:45-50,52-67). */
const char *thmm_last_error(void);

/* ------------------------------------------------------------------------ */
/* Rake-compress plan                                                       */
/* ------------------------------------------------------------------------ */

/* The op records carry the field meaning of the planner structs the
 * reference consumes (btrc::Rake, BranchCombination, BranchAbsorption,
 * Compression; used at src/inference.cc:155-214,363-412 and sized at
 * src/gpu_backend_impl.inc:39-42). Edge fields name edge slots in the
 * reference convention: a compression overwrites its left_edge slot with the
 * composed parent->child path. */
typedef struct thmm_rake {
  uint32_t leaf, parent, edge, branch;
} thmm_rake;
typedef struct thmm_branch_combination {
  uint32_t destination, source, stage;
} thmm_branch_combination;
typedef struct thmm_branch_absorption {
  uint32_t parent, branch, round;
} thmm_branch_absorption;
typedef struct thmm_compression {
  uint32_t left_edge, middle, right_edge, parent, child, tape;
} thmm_compression;

typedef enum thmm_primitive {
  THMM_RAKE = 0,
  THMM_BRANCH_COMBINATION = 1,
  THMM_BRANCH_ABSORPTION = 2,
  THMM_COMPRESSION = 3
} thmm_primitive;

/* One level of the contraction: `count` write-disjoint ops of one kind,
 * starting at `offset` in that kind's op array (btrc::PrimitiveBatch). */
typedef struct thmm_primitive_batch {
  uint32_t primitive, offset, count;
} thmm_primitive_batch;

/* Flat view of a plan. Arrays are owned by the plan object. */
typedef struct thmm_plan_arrays {
  size_t num_nodes, num_edges, num_branches, num_compressions, num_roots;
  size_t num_rakes, num_combinations, num_absorptions, num_batches;
  size_t num_rounds;
  const uint32_t *roots;         /* [num_roots], ascending */
  const uint32_t *edge_parents;  /* [num_edges] */
  const uint32_t *edge_children; /* [num_edges] */
  const thmm_rake *rakes;
  const thmm_branch_combination *combinations;
  const thmm_branch_absorption *absorptions;
  const thmm_compression *compressions; /* [num_compressions] */
  const thmm_primitive_batch *batches;
  /* Per round: index of the round's first batch, [num_rounds + 1]. */
  const uint32_t *round_batch_offsets;
} thmm_plan_arrays;

typedef struct thmm_plan thmm_plan;

/* Builds a rake-compress plan from a parent array (parents[i] == -1 marks a
 * root; several roots make a forest whose trees are contracted together).
 * Replaces btrc::MakePlan(std::vector<int64_t>) as called at
 * tests/inference_test.cc:159-160 and tests/accelerator_test.h:19-20.
 * A forest of at least 65,536 nodes whose trees each occupy the node range
 * from their root to the next root is planned tree by tree on host threads
 * (environment variable THMM_PLAN_THREADS, default the hardware threads up
 * to 16; 1 plans sequentially); the plan is the same either way. */
thmm_status thmm_plan_create(const int64_t *parents, size_t num_nodes,
                             thmm_plan **out);

/* The same plan as thmm_plan_create (identical arrays), built on HIP device
 * `device`: per round, leaves are compacted, branches sorted by parent and
 * unary chains ranked by pointer jumping on the device, and only the ops
 * come back to the host. `parents` is a host array, or with
 * parents_on_device != 0 a device array on that device. Same errors as
 * thmm_plan_create; at most 2^31 - 1 nodes (length_error). No reference
 * counterpart: the reference plans on the host through btrc::MakePlan
 * (tests/accelerator_test.h:19-20). */
thmm_status thmm_plan_create_device(const int64_t *parents, size_t num_nodes,
                                    int parents_on_device, int device,
                                    thmm_plan **out);

/* Wraps an externally built plan (e.g. a btrc::Plan flattened by the C++
 * layer). The arrays are copied and checked: every non-root node must be
 * eliminated exactly once and every batch must be write-disjoint. */
thmm_status thmm_plan_create_from_arrays(const thmm_plan_arrays *arrays,
                                         thmm_plan **out);

thmm_status thmm_plan_get_arrays(const thmm_plan *plan, thmm_plan_arrays *out);
void thmm_plan_destroy(thmm_plan *plan);

/* ------------------------------------------------------------------------ */
/* Device + workspace                                                       */
/* ------------------------------------------------------------------------ */

/* 1 if a HIP device is present (src/gpu_backend_impl.inc:1782-1785). */
int thmm_available(void);
/* "<name> (<gcnArchName>)" (src/gpu_backend_impl.inc:1787-1799). */
thmm_status thmm_device_description(int device, char *buffer, size_t size);

typedef enum thmm_precision {
  THMM_F32 = 0, /* Scalar = float: FP32 values, FP64 running log scales */
  THMM_F64 = 1  /* Scalar = double: plain FP64 everywhere */
} thmm_precision;

enum {
  THMM_RESERVE_MAXIMUM = 1u,   /* Workspace::ReserveMaximum */
  THMM_RESERVE_SAMPLING = 2u,  /* Workspace::ReserveSampling */
  THMM_RESERVE_MARGINALS = 4u, /* Workspace::ReserveMarginals */
  /* No reference counterpart: allocate no pinned staging and no
   * workspace-owned outputs; only the *_device entry points may be used.
   * Keeps multi-GB edge-marginal outputs single-copy in HBM. */
  THMM_RESERVE_DEVICE_RESIDENT = 8u,
  /* No reference counterpart: the device-resident entry points replay a
   * hipGraph of their whole launch sequence, captured at the first call and
   * again whenever the pointers, batch or reduction flag change; the replay
   * is enqueued on the call's stream. thmm_profile_marginals_device stays
   * eager. */
  THMM_RESERVE_GRAPH = 16u,
  /* No reference counterpart: thmm_marginals_device (and its profiling and
   * autotuning forms) and thmm_logz_backward_device write edge outputs
   * [B][E][K][K] as IEEE binary16 (2 bytes a value, round to nearest even;
   * magnitudes below 2^-14 are subnormal, absolute resolution 2^-24; finite
   * up to 65504) instead of the reservation's precision, halving the
   * largest output stream. Node marginals, log partitions and all internal
   * state keep the reservation's precision. Needs THMM_RESERVE_MARGINALS |
   * THMM_RESERVE_DEVICE_RESIDENT; the edge output must be 2-byte aligned
   * (16-byte aligned for the vectorised kernels); batch-reduced edge
   * gradients are not available with it. */
  THMM_RESERVE_HALF_EDGE_MARGINALS = 32u
};

/* tree_hmm::CategoricalInputUpdate (include/tree_hmm/accelerator.h:92-96). */
typedef enum thmm_categorical_update {
  THMM_UPDATE_ALL = 0,
  THMM_UPDATE_FACTORS = 1,
  THMM_UPDATE_NONE = 2
} thmm_categorical_update;

/* tree_hmm::AcceleratorTimings (include/tree_hmm/accelerator.h:98-111). */
typedef struct thmm_timings {
  double upload_ms, kernel_ms, download_ms, wall_ms;
} thmm_timings;

typedef struct thmm_workspace thmm_workspace;

thmm_status thmm_workspace_create(thmm_workspace **out);
void thmm_workspace_destroy(thmm_workspace *workspace);

/* Workspace::Reserve / ReserveMaximum / ReserveSampling / ReserveMarginals
 * (include/tree_hmm/detail/gpu_backend_api.inc:19-51 and
 * src/gpu_backend_impl.inc:1803-2076). `identity` is the address the caller
 * uses to name its plan object; later calls must present the same identity
 * (src/gpu_backend_impl.inc:2162-2166). The plan must outlive the
 * reservation. `flags` is a bitwise OR of THMM_RESERVE_*. */
thmm_status thmm_reserve(thmm_workspace *workspace, const thmm_plan *plan,
                         const void *identity, size_t states,
                         size_t batch_capacity, thmm_precision precision,
                         unsigned flags, int device);

/* Workspace::ReserveCategorical{,Maximum,Sampling,Marginals}
 * (src/gpu_backend_impl.inc:1981-2031). */
thmm_status thmm_reserve_categorical(
    thmm_workspace *workspace, const thmm_plan *plan, const void *identity,
    size_t states, size_t batch_capacity, size_t categories,
    const uint32_t *observation_nodes, size_t num_observation_nodes,
    thmm_precision precision, unsigned flags, int device);

/* Workspace::Inputs(batch): pinned host storage laid out as
 * BatchedModelView ([batch, node, state] and [edge, parent, child]).
 * (src/gpu_backend_impl.inc:2078-2102). Element type follows the
 * workspace precision. */
thmm_status thmm_inputs(thmm_workspace *workspace, size_t batch,
                        void **node_potentials, size_t *node_count,
                        void **edge_potentials, size_t *edge_count);

/* Workspace::CategoricalInputs(batch) (src/gpu_backend_impl.inc:2104-2134). */
thmm_status thmm_categorical_inputs(thmm_workspace *workspace, size_t batch,
                                    uint8_t **observations,
                                    size_t *observation_count,
                                    void **root_potential,
                                    void **emission_potentials,
                                    size_t *emission_count,
                                    void **edge_potentials, size_t *edge_count);

/* Workspace::Uniforms(batch) (src/gpu_backend_impl.inc:2136-2150). */
thmm_status thmm_uniforms(thmm_workspace *workspace, size_t batch,
                          void **uniforms, size_t *count);

/* BatchedModelView (include/tree_hmm/accelerator.h:15-23). */
typedef struct thmm_dense_model {
  const void *identity;
  size_t states, batch;
  const void *node_potentials; /* [batch, node, state] */
  size_t node_count;
  const void *edge_potentials; /* [edge, parent state, child state] */
  size_t edge_count;
} thmm_dense_model;

/* BatchedCategoricalModelView (include/tree_hmm/accelerator.h:45-60). */
typedef struct thmm_categorical_model {
  const void *identity;
  size_t states, batch, categories;
  const uint32_t *observation_nodes;
  size_t num_observation_nodes;
  const uint8_t *observations; /* [batch, observation node] */
  size_t observation_count;
  const void *root_potential; /* [state] */
  size_t root_count;
  const void *emission_potentials; /* [category, state] */
  size_t emission_count;
  const void *edge_potentials; /* [edge, parent state, child state] */
  size_t edge_count;
} thmm_categorical_model;

/* Result spans point into workspace-owned pinned memory and stay valid until
 * the next call, Reserve or destruction (include/tree_hmm/accelerator.h:
 * 113-146). Value element types follow the workspace precision. Partition
 * outputs hold batch * num_roots values, batch-major. */

/* PartitionFunctionPrepared (log_domain = 0) and
 * LogPartitionFunctionPrepared (log_domain = 1)
 * (src/gpu_backend_impl.inc:3021-3031). */
thmm_status thmm_partition(thmm_workspace *workspace,
                           const thmm_dense_model *model, int log_domain,
                           const void **values, size_t *count,
                           thmm_timings *timings);
/* Categorical overloads with CategoricalInputUpdate
 * (src/gpu_backend_impl.inc:3033-3059). */
thmm_status thmm_partition_categorical(thmm_workspace *workspace,
                                       const thmm_categorical_model *model,
                                       thmm_categorical_update update,
                                       int log_domain, const void **values,
                                       size_t *count, thmm_timings *timings);

/* MaximumAPosterioriPrepared (src/gpu_backend_impl.inc:3061-3079). */
thmm_status thmm_maximum(thmm_workspace *workspace,
                         const thmm_dense_model *model,
                         const void **log_weights, const uint32_t **states,
                         thmm_timings *timings);
thmm_status thmm_maximum_categorical(thmm_workspace *workspace,
                                     const thmm_categorical_model *model,
                                     thmm_categorical_update update,
                                     const void **log_weights,
                                     const uint32_t **states,
                                     thmm_timings *timings);

/* PosteriorSamplePrepared (src/gpu_backend_impl.inc:3081-3110). `uniforms`
 * holds one U[0,1) variate per batch item and node, [batch, node]. */
thmm_status thmm_sample(thmm_workspace *workspace,
                        const thmm_dense_model *model, const void *uniforms,
                        size_t uniform_count, const uint32_t **states,
                        thmm_timings *timings);
thmm_status thmm_sample_categorical(thmm_workspace *workspace,
                                    const thmm_categorical_model *model,
                                    const void *uniforms, size_t uniform_count,
                                    thmm_categorical_update update,
                                    const uint32_t **states,
                                    thmm_timings *timings);

/* PosteriorMarginalsPrepared (src/gpu_backend_impl.inc:3112-3130): log Z
 * [batch * roots], node marginals [batch, node, state] and edge marginals
 * [batch, edge, parent state, child state]. */
thmm_status thmm_marginals(thmm_workspace *workspace,
                           const thmm_dense_model *model,
                           const void **log_partitions, const void **nodes,
                           const void **edges, thmm_timings *timings);
thmm_status thmm_marginals_categorical(thmm_workspace *workspace,
                                       const thmm_categorical_model *model,
                                       thmm_categorical_update update,
                                       const void **log_partitions,
                                       const void **nodes, const void **edges,
                                       thmm_timings *timings);

/* The FP64 log partitions of the latest partition/marginal call, before
 * rounding to the workspace Scalar ([batch * roots]). No reference
 * counterpart: the reference returns Scalar only. */
thmm_status thmm_result_log_partitions_f64(thmm_workspace *workspace,
                                           const double **values,
                                           size_t *count);

/* ------------------------------------------------------------------------ */
/* Branch-length gradient pair (categorical models)                         */
/* ------------------------------------------------------------------------ */
/* No reference counterpart. For edge factors P_e = expm(Q t_e) staged by the
 * caller (e.g. through thmm_categorical_inputs with THMM_UPDATE_FACTORS
 * reuse of the observations), the forward half returns the FP64 log
 * partitions [batch * roots] and keeps the contraction state; the backward
 * half runs the adjoint pass seeded with grad_log_z (host, FP64
 * [batch * roots]; NULL = all ones) and returns
 *   grad[e] = d (sum_b grad_log_z[b] log Z_b) / d t_e
 *           = sum_{p,c} M_e[p][c] (Q P_e)[p][c] / P_e[p][c]
 * (host, FP64 [num_edges]), M_e the batch-summed edge marginals, Q the
 * row-major K x K rate matrix (host, FP64). Entries with P_e[p][c] = 0 add
 * nothing, so the chain needs P_e > 0, which expm(Q t) has for t > 0 and an
 * irreducible Q. Needs a ReserveCategoricalMarginals workspace; any other
 * call in between invalidates the forward state (logic_error). Result spans
 * live in workspace pinned memory until the next call. */
thmm_status thmm_categorical_logz_forward(thmm_workspace *workspace,
                                          const thmm_categorical_model *model,
                                          thmm_categorical_update update,
                                          const double **log_z, size_t *count);
/* The forward half with the batch (sites) summed on the device: sums[r] =
 * sum over b of log Z_{b,r} (FP64 [roots]; 256 partial sums, partial t
 * adding b = t, t + 256, ... in turn, then added pairwise: t and t + 128,
 * then t and t + 64, ..., t and t + 1), the log likelihood of the whole
 * alignment for each tree; only these roots values
 * leave the device. Holds the contraction state for
 * thmm_branch_length_backward like thmm_categorical_logz_forward (a NULL
 * seed there then gives the gradient of the summed log likelihood). Same
 * errors as thmm_categorical_logz_forward (domain_error when a site's
 * partition function is not positive). No reference counterpart: the
 * reference returns per-item values and leaves the sum to the caller
 * (src/gpu_backend_impl.inc:2917-2941 writes [B] results). */
thmm_status thmm_categorical_log_likelihood(thmm_workspace *workspace,
                                            const thmm_categorical_model *model,
                                            thmm_categorical_update update,
                                            const double **sums, size_t *count);
thmm_status thmm_branch_length_backward(thmm_workspace *workspace,
                                        const double *grad_log_z,
                                        const double *rate_matrix,
                                        const double **grad, size_t *count);

/* No reference counterpart (the reference takes edge factors only as
 * caller-staged values, include/tree_hmm/accelerator.h:86-96). Transition
 * matrices of a rate matrix Q (host, FP64, row-major K x K) reversible with
 * respect to the stationary law pi (host, FP64 [K], positive; detailed
 * balance pi_i Q_ij = pi_j Q_ji within 1e-10 of the largest flux):
 *   P_e = expm(Q t_e) = V diag(exp(lambda t_e)) V^-1,
 * from the symmetric eigendecomposition of D^1/2 Q D^-1/2 (D = diag(pi)),
 * for branch lengths t (host, FP64 [edges], finite, >= 0). Entries that
 * round below zero are set to zero. invalid_argument otherwise.
 *
 * thmm_expm_reversible writes the matrices to out (host, FP64
 * [edges][K][K]) on the host. thmm_categorical_set_branch_lengths forms
 * them on the device, rounded to the workspace precision, as the edge
 * factors of a categorical workspace (one length per plan edge, in edge
 * order). The observations and the root and emission factors must already
 * be resident (a call with THMM_UPDATE_ALL or THMM_UPDATE_FACTORS;
 * logic_error otherwise); later calls with THMM_UPDATE_NONE use the new
 * edges, THMM_UPDATE_FACTORS restages the host edges. It drops a held
 * forward pass of the branch-length pair and blocks until the factors are
 * written. */
thmm_status thmm_expm_reversible(const double *rate_matrix,
                                 const double *stationary, size_t states,
                                 const double *branch_lengths, size_t edges,
                                 double *out);
thmm_status thmm_categorical_set_branch_lengths(thmm_workspace *workspace,
                                                const double *rate_matrix,
                                                const double *stationary,
                                                const double *branch_lengths,
                                                size_t edge_count);

/* ------------------------------------------------------------------------ */
/* Device-resident entry points                                             */
/* ------------------------------------------------------------------------ */
/* Inputs already in HBM, outputs written to caller device buffers, work
 * enqueued on `stream` without a host synchronisation: a hipStream_t, NULL
 * for the workspace's own stream, or THMM_HIP_NULL_STREAM for HIP's null
 * (default) stream, whose handle is itself 0. Element type follows the
 * workspace precision.
 * These have no reference counterpart; they serve callers that keep
 * potentials resident (the reference uploads every call,
 * src/gpu_backend_impl.inc:2197-2213). */

#define THMM_HIP_NULL_STREAM ((void *)(~(uintptr_t)0))

/* Linear potentials (these two entry points) are taken as given: node
 * vectors are moved to the log domain with an FP64 scale, but edge matrices
 * enter the products unscaled, so with THMM_F32 a product of two edge
 * entries (and K times it) must stay below the FP32 maximum (entries up to
 * about 1e17 at K = 64). Wider ranges belong in the log-potential pair
 * below, which normalises every edge row by its maximum first, or in
 * THMM_F64. */

/* Scaled log partitions: d_log_z is double [batch * roots]. */
thmm_status thmm_log_partition_device(thmm_workspace *workspace,
                                      const void *d_node_potentials,
                                      const void *d_edge_potentials,
                                      size_t batch, double *d_log_z,
                                      void *stream);

/* log Z + adjoint: node marginals [batch, node, state], edge marginals
 * [batch, edge, parent, child] (= d log Z / d log potentials). */
thmm_status thmm_marginals_device(thmm_workspace *workspace,
                                  const void *d_node_potentials,
                                  const void *d_edge_potentials, size_t batch,
                                  double *d_log_z, void *d_node_marginals,
                                  void *d_edge_marginals, void *stream);

/* Forward half of the adjoint pair: log potentials in, log Z out. Keeps the
 * contraction state in the workspace for thmm_logz_backward_device. */
thmm_status thmm_logz_forward_device(thmm_workspace *workspace,
                                     const void *d_log_node_potentials,
                                     const void *d_log_edge_potentials,
                                     size_t batch, double *d_log_z,
                                     void *stream);

/* Backward half: d_grad_log_z (double [batch * roots], NULL = all ones)
 * scales the adjoint. d_grad_nodes is [batch, node, state]. d_grad_edges is
 * [batch, edge, parent, child], or [edge, parent, child] summed over the
 * batch when reduce_edges_over_batch != 0 (the edges are broadcast). */
thmm_status thmm_logz_backward_device(thmm_workspace *workspace,
                                      const double *d_grad_log_z,
                                      void *d_grad_nodes, void *d_grad_edges,
                                      int reduce_edges_over_batch,
                                      void *stream);

/* thmm_maximum on device-resident dense potentials (a THMM_RESERVE_MAXIMUM
 * reservation, device-resident or not): the same kernels and outputs as
 * RunMaximumPrepared (src/gpu_backend_impl.inc:2405-2604), written to
 * d_log_weights (double
 * [batch * roots], -inf where no assignment has positive weight, which the
 * host call reports as an error) and d_states (uint32 [batch, node]). */
thmm_status thmm_maximum_device(thmm_workspace *workspace,
                                const void *d_node_potentials,
                                const void *d_edge_potentials, size_t batch,
                                double *d_log_weights, uint32_t *d_states,
                                void *stream);

/* thmm_sample on device-resident dense potentials and variates (a
 * THMM_RESERVE_SAMPLING reservation): the same kernels and draws as the
 * sampling mode of RunPrepared (src/gpu_backend_impl.inc:2156-2404, draws
 * at 2328-2384) from d_uniforms [batch, node] in the
 * workspace precision. The variates are not checked on the host; any value
 * gives in-range states (one at or above 1 takes the last supported state).
 * d_log_z (double [batch * roots]) is -inf for a nonpositive partition
 * function, and *d_failure (uint32, zeroed by the call) is set to 1 when a
 * conditional has no support; the host call reports both as errors. */
thmm_status thmm_sample_device(thmm_workspace *workspace,
                               const void *d_node_potentials,
                               const void *d_edge_potentials,
                               const void *d_uniforms, size_t batch,
                               double *d_log_z, uint32_t *d_states,
                               uint32_t *d_failure, void *stream);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc). thmm_log_partition_device without a workspace: the same
 * inputs ([batch, node, state] node and [edge, parent, child] edge
 * potentials in HBM, precision's element type), the same output (d_log_z,
 * double [batch * roots]) from the plan's levels, with every buffer beyond
 * the inputs and outputs (the renamed ops, node, branch and path state) in
 * one library-owned static device array of THMM_STATIC_SCRATCH_BYTES per
 * GPU, which that GPU holds until the process ends. The call allocates no
 * device memory; thmm_static_scratch_bytes gives the bytes a shape needs,
 * and a shape above the maximum is a length_error. The topology is copied
 * from the plan (host memory) on `stream` (NULL or THMM_HIP_NULL_STREAM:
 * HIP's null stream); `device` is made current. Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other
 * threads) race, so serialise them. */
#define THMM_STATIC_SCRATCH_BYTES ((size_t)64 << 20)
thmm_status thmm_static_scratch_bytes(const thmm_plan *plan, size_t states, size_t batch,
                                      thmm_precision precision, size_t *bytes);
thmm_status thmm_log_partition_device_static(const thmm_plan *plan, size_t states, size_t batch,
                                             thmm_precision precision, const void *d_node_potentials,
                                             const void *d_edge_potentials, double *d_log_z, void *stream,
                                             int device);

/* Blocks until the workspace stream has drained. */
thmm_status thmm_synchronize(thmm_workspace *workspace);

/* No reference counterpart (the reference stores every tape and path,
 * src/gpu_backend_impl.inc:1913-1966). A budget, in bytes, for the path
 * storage of the workspace's next reservations: the composed paths, their
 * FP64 row scales and, for marginal reservations, their adjoints (0, the
 * default: no budget). When a reservation's path storage would exceed it,
 * a batch-1 forest whose trees occupy consecutive node ranges (device-
 * resident, without MAP or sampling) runs in groups of consecutive trees
 * that reuse one path store, each group's storage within the budget;
 * anything else is a length_error, and so is a single tree above the
 * budget. In tree groups thmm_marginals_device (and thmm_log_partition_device)
 * runs each group's forward and reverse pass in turn, and
 * thmm_logz_backward_device recomputes each group's paths from the stored
 * node state and the caller's log edges (which must stay unchanged between
 * the forward and the backward call) before the group's reverse pass; the
 * results are bit-equal to an unbudgeted workspace's. thmm_path_storage
 * reports the groups of the current reservation (0: not grouped) and the
 * bytes of path storage it holds. */
thmm_status thmm_set_path_budget(thmm_workspace *workspace, size_t bytes);
thmm_status thmm_path_storage(const thmm_workspace *workspace, size_t *groups, size_t *bytes);

/* ------------------------------------------------------------------------ */
/* Launch tuning                                                            */
/* ------------------------------------------------------------------------ */
/* No reference counterpart. Launch variants of a reserved workspace:
 * compress_items_forward / _reverse: K = 16 FP32 MFMA compressions per
 * wavefront (1 or 2); rake_block: threads per block of the team rake kernels
 * (64, 128 or 256); gather_chunk: rakes a team folds into one parent per
 * step of the fused forward pass (1..64); compress_depth: compression-tree
 * levels evaluated as one grouped product of up to 2^depth slots on the
 * K = 16 FP32 path (1 = every compression stored, 2 or 3; default 2); in
 * MAP calls at K = 8, 16 and 32, depth 2 or 3 evaluates each compression
 * together with the input compressions absorbed into it (paths kept in
 * registers, the same values and choices) and 1 runs every compression
 * level on its own;
 * rake_log_edges: in thmm_logz_forward_device / _backward_device with
 * K = 8, 16, 32 or 64, the rake kernels read the log edge potentials in
 * place (1, default; at K = 8 and 16 and in FP64 only the edges
 * compressions read are staged, at K = 32 and 64 in FP32 the compressions
 * read theirs in place too; the caller's log edges must stay unchanged
 * until the backward call) or every edge is staged first (0); both give the
 * same values; persistent: 0 (every level is its own launch). The value 1
 * asks for each pass as one persistent launch whose blocks pull level
 * chunks from a device queue; that mode is switched off in this build (its
 * launches did not complete on the MI355X) and is refused with logic_error;
 * input_leaves: the rakes of original leaves form their leaf vectors from
 * the call's inputs inside the rake kernels, and the node initialisation
 * skips those leaves (1, default), or every node's state is written first
 * (0). It covers categorical partition, marginal and branch-length calls
 * whose rake levels run on the generic kernels (K other than 8, 16, 32 and
 * 64, no fused forward program: the observations, emissions and root
 * potential), and dense partition and marginal calls on linear potentials
 * with the fused forward program (K = 8, 16, 32, 64; 16-byte aligned
 * buffers: round 0's rakes read the caller's potentials, and so do round
 * 0's absorptions for their parents, which the initialisation also skips).
 * Dense MAP calls at K = 8, 16 and 32 (16-byte aligned node potentials)
 * likewise log round 0's leaves and parents where they first read them
 * (1), or every node first (0). Both settings
 * give the same values bit for bit; a change drops a held forward pass.
 * Values outside these sets are invalid_argument. Setting them drops
 * captured graphs; a new gather_chunk or compress_depth rebuilds the fused
 * forward program, and a new compress_depth also drops a held
 * thmm_logz_forward_device pass. ms is filled by the autotuner. */
typedef struct thmm_tuning {
  uint32_t compress_items_forward, compress_items_reverse;
  uint32_t rake_block, gather_chunk;
  uint32_t compress_depth;
  uint32_t rake_log_edges;
  uint32_t persistent;
  uint32_t input_leaves;
  double ms;
} thmm_tuning;

thmm_status thmm_get_tuning(const thmm_workspace *workspace, thmm_tuning *out);
thmm_status thmm_set_tuning(thmm_workspace *workspace, const thmm_tuning *tuning);

/* Autotuner: runs thmm_marginals_device on the given inputs and outputs once
 * per candidate launch variant (coordinate descent over the four knobs,
 * each candidate the best of three event-timed runs), keeps the fastest
 * setting in the workspace and reports it, with its milliseconds per call,
 * in *chosen. Blocks until done; the outputs hold valid marginals. */
thmm_status thmm_autotune_marginals_device(
    thmm_workspace *workspace, const void *d_node_potentials,
    const void *d_edge_potentials, size_t batch, double *d_log_z,
    void *d_node_marginals, void *d_edge_marginals, void *stream,
    thmm_tuning *chosen);

/* ------------------------------------------------------------------------ */
/* Measurement                                                              */
/* ------------------------------------------------------------------------ */

/* One launch of a profiled call. phase: 0 staging, 1 forward level,
 * 2 finish + adjoint seed, 3 reverse level, 4 output transpose. primitive
 * is a thmm_primitive for phases 1 and 3, or, in phase 1, 4 for a fused
 * rake level (rakes, combinations and absorptions of a round in one launch;
 * count = rakes) and 5 for its partial-absorption level (count = partials);
 * in phases 1 and 3, 6 for a level of grouped compressions (count = K-element
 * vectors read and written per batch item: (n + 1) K + n - 1 forward and
 * (2n + 1) K + 2 (n - 1) reverse for a group of n slots), and in phase 3, 7
 * for the part of such a level whose groups have 3 or 4 slots and 8 for the
 * part whose groups have 2; otherwise count is the level's op count. ops is
 * the launch's number of work units (groups, for grouped compressions). */
typedef struct thmm_launch_timing {
  uint32_t phase, primitive, count, batch;
  double ms;
  /* Work units of the launch: rakes, partial sums, compressions or
   * compression groups (equal to count except for grouped compressions). */
  uint32_t ops;
} thmm_launch_timing;

/* thmm_marginals_device with a HIP event pair recorded on the launch stream
 * around every kernel launch; blocks until done and reports up to
 * `capacity` launch durations. No reference counterpart: it replaces the
 * coarse AcceleratorTimings::kernel_ms with per-level figures. */
thmm_status thmm_profile_marginals_device(
    thmm_workspace *workspace, const void *d_node_potentials,
    const void *d_edge_potentials, size_t batch, double *d_log_z,
    void *d_node_marginals, void *d_edge_marginals, void *stream,
    thmm_launch_timing *timings, size_t capacity, size_t *count);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* TREE_HMM_AMD_H_ */
