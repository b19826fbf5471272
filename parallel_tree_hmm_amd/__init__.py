"""MI355X-native batched tree-HMM inference (drop-in for the tree_hmm::rocm
accelerator path of joaospinto/parallel_tree_hmm).

The compute runs in hand-written gfx950 HIP kernels inside
libtree_hmm_amd.so, reached through its C-ABI (include/tree_hmm_amd.h). This
package is the Python mirror of the reference accelerator API; it raises if
the library has not been built.
"""

from .api import (  # noqa: F401
    AcceleratorRuntimeError,
    AcceleratorTimings,
    BatchedCategoricalModelView,
    BatchedMarginalView,
    BatchedMaximumAssignmentView,
    BatchedModelView,
    BatchedPosteriorSampleView,
    CategoricalInputUpdate,
    DomainError,
    InvalidArgument,
    LengthError,
    LogicError,
    MutableBatchedCategoricalModelView,
    MutableBatchedModelView,
    PartitionView,
    Plan,
    TreeHmmError,
    Workspace,
    available,
    device_description,
    expm_reversible,
    log_partition_device_static,
    log_partition_function_prepared,
    make_plan,
    maximum_a_posteriori_prepared,
    partition_function_prepared,
    posterior_marginals_prepared,
    posterior_sample_prepared,
    static_scratch_bytes,
)
from ._abi import library_path  # noqa: F401
