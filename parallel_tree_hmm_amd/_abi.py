"""ctypes declarations of the C-ABI in include/tree_hmm_amd.h.

The library is loaded once per process. torch (when importable) is imported
first so that torch's HIP runtime, which carries the same soname
(libamdhip64.so.7) as /opt/rocm's, is the single runtime in the process and
device pointers from torch tensors are valid in the library.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE / "libtree_hmm_amd.so"

c_size = ctypes.c_size_t
c_u32 = ctypes.c_uint32
c_int = ctypes.c_int
c_void_p = ctypes.c_void_p
c_status = ctypes.c_int

THMM_OK = 0
THMM_INVALID_ARGUMENT = 1
THMM_LOGIC_ERROR = 2
THMM_LENGTH_ERROR = 3
THMM_DOMAIN_ERROR = 4
THMM_RUNTIME_ERROR = 5

THMM_F32 = 0
THMM_F64 = 1

THMM_RESERVE_MAXIMUM = 1
THMM_RESERVE_SAMPLING = 2
THMM_RESERVE_MARGINALS = 4
THMM_RESERVE_DEVICE_RESIDENT = 8
THMM_RESERVE_GRAPH = 16
THMM_RESERVE_HALF_EDGE_MARGINALS = 32
THMM_STATIC_SCRATCH_BYTES = 64 << 20

THMM_RAKE = 0
THMM_BRANCH_COMBINATION = 1
THMM_BRANCH_ABSORPTION = 2
THMM_COMPRESSION = 3


class Rake(ctypes.Structure):
    _fields_ = [("leaf", c_u32), ("parent", c_u32), ("edge", c_u32), ("branch", c_u32)]


class BranchCombination(ctypes.Structure):
    _fields_ = [("destination", c_u32), ("source", c_u32), ("stage", c_u32)]


class BranchAbsorption(ctypes.Structure):
    _fields_ = [("parent", c_u32), ("branch", c_u32), ("round", c_u32)]


class Compression(ctypes.Structure):
    _fields_ = [
        ("left_edge", c_u32),
        ("middle", c_u32),
        ("right_edge", c_u32),
        ("parent", c_u32),
        ("child", c_u32),
        ("tape", c_u32),
    ]


class PrimitiveBatch(ctypes.Structure):
    _fields_ = [("primitive", c_u32), ("offset", c_u32), ("count", c_u32)]


class PlanArrays(ctypes.Structure):
    _fields_ = [
        ("num_nodes", c_size),
        ("num_edges", c_size),
        ("num_branches", c_size),
        ("num_compressions", c_size),
        ("num_roots", c_size),
        ("num_rakes", c_size),
        ("num_combinations", c_size),
        ("num_absorptions", c_size),
        ("num_batches", c_size),
        ("num_rounds", c_size),
        ("roots", ctypes.POINTER(c_u32)),
        ("edge_parents", ctypes.POINTER(c_u32)),
        ("edge_children", ctypes.POINTER(c_u32)),
        ("rakes", ctypes.POINTER(Rake)),
        ("combinations", ctypes.POINTER(BranchCombination)),
        ("absorptions", ctypes.POINTER(BranchAbsorption)),
        ("compressions", ctypes.POINTER(Compression)),
        ("batches", ctypes.POINTER(PrimitiveBatch)),
        ("round_batch_offsets", ctypes.POINTER(c_u32)),
    ]


class Timings(ctypes.Structure):
    _fields_ = [
        ("upload_ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("download_ms", ctypes.c_double),
        ("wall_ms", ctypes.c_double),
    ]


class LaunchTiming(ctypes.Structure):
    _fields_ = [
        ("phase", c_u32),
        ("primitive", c_u32),
        ("count", c_u32),
        ("batch", c_u32),
        ("ms", ctypes.c_double),
        ("ops", c_u32),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("compress_items_forward", c_u32),
        ("compress_items_reverse", c_u32),
        ("rake_block", c_u32),
        ("gather_chunk", c_u32),
        ("compress_depth", c_u32),
        ("rake_log_edges", c_u32),
        ("persistent", c_u32),
        ("input_leaves", c_u32),
        ("ms", ctypes.c_double),
    ]


class DenseModel(ctypes.Structure):
    _fields_ = [
        ("identity", c_void_p),
        ("states", c_size),
        ("batch", c_size),
        ("node_potentials", c_void_p),
        ("node_count", c_size),
        ("edge_potentials", c_void_p),
        ("edge_count", c_size),
    ]


class CategoricalModel(ctypes.Structure):
    _fields_ = [
        ("identity", c_void_p),
        ("states", c_size),
        ("batch", c_size),
        ("categories", c_size),
        ("observation_nodes", c_void_p),
        ("num_observation_nodes", c_size),
        ("observations", c_void_p),
        ("observation_count", c_size),
        ("root_potential", c_void_p),
        ("root_count", c_size),
        ("emission_potentials", c_void_p),
        ("emission_count", c_size),
        ("edge_potentials", c_void_p),
        ("edge_count", c_size),
    ]


P = ctypes.POINTER
PP = lambda t: ctypes.POINTER(ctypes.POINTER(t))  # noqa: E731

# name -> (restype, argtypes)
PROTOTYPES = {
    "thmm_last_error": (ctypes.c_char_p, []),
    "thmm_plan_create": (c_status, [P(ctypes.c_int64), c_size, P(c_void_p)]),
    "thmm_plan_create_device": (c_status, [c_void_p, c_size, c_int, c_int, P(c_void_p)]),
    "thmm_plan_create_from_arrays": (c_status, [P(PlanArrays), P(c_void_p)]),
    "thmm_plan_get_arrays": (c_status, [c_void_p, P(PlanArrays)]),
    "thmm_plan_destroy": (None, [c_void_p]),
    "thmm_available": (c_int, []),
    "thmm_device_description": (c_status, [c_int, ctypes.c_char_p, c_size]),
    "thmm_workspace_create": (c_status, [P(c_void_p)]),
    "thmm_workspace_destroy": (None, [c_void_p]),
    "thmm_reserve": (c_status, [c_void_p, c_void_p, c_void_p, c_size, c_size, c_int, ctypes.c_uint, c_int]),
    "thmm_reserve_categorical": (
        c_status,
        [c_void_p, c_void_p, c_void_p, c_size, c_size, c_size, c_void_p, c_size, c_int, ctypes.c_uint, c_int],
    ),
    "thmm_inputs": (c_status, [c_void_p, c_size, P(c_void_p), P(c_size), P(c_void_p), P(c_size)]),
    "thmm_categorical_inputs": (
        c_status,
        [c_void_p, c_size, P(c_void_p), P(c_size), P(c_void_p), P(c_void_p), P(c_size), P(c_void_p), P(c_size)],
    ),
    "thmm_uniforms": (c_status, [c_void_p, c_size, P(c_void_p), P(c_size)]),
    "thmm_partition": (c_status, [c_void_p, P(DenseModel), c_int, P(c_void_p), P(c_size), P(Timings)]),
    "thmm_partition_categorical": (
        c_status,
        [c_void_p, P(CategoricalModel), c_int, c_int, P(c_void_p), P(c_size), P(Timings)],
    ),
    "thmm_maximum": (c_status, [c_void_p, P(DenseModel), P(c_void_p), P(c_void_p), P(Timings)]),
    "thmm_maximum_categorical": (
        c_status,
        [c_void_p, P(CategoricalModel), c_int, P(c_void_p), P(c_void_p), P(Timings)],
    ),
    "thmm_sample": (c_status, [c_void_p, P(DenseModel), c_void_p, c_size, P(c_void_p), P(Timings)]),
    "thmm_sample_categorical": (
        c_status,
        [c_void_p, P(CategoricalModel), c_void_p, c_size, c_int, P(c_void_p), P(Timings)],
    ),
    "thmm_marginals": (
        c_status,
        [c_void_p, P(DenseModel), P(c_void_p), P(c_void_p), P(c_void_p), P(Timings)],
    ),
    "thmm_marginals_categorical": (
        c_status,
        [c_void_p, P(CategoricalModel), c_int, P(c_void_p), P(c_void_p), P(c_void_p), P(Timings)],
    ),
    "thmm_result_log_partitions_f64": (c_status, [c_void_p, P(c_void_p), P(c_size)]),
    "thmm_categorical_logz_forward": (
        c_status,
        [c_void_p, P(CategoricalModel), c_int, P(c_void_p), P(c_size)],
    ),
    "thmm_categorical_log_likelihood": (
        c_status,
        [c_void_p, P(CategoricalModel), c_int, P(c_void_p), P(c_size)],
    ),
    "thmm_branch_length_backward": (c_status, [c_void_p, c_void_p, c_void_p, P(c_void_p), P(c_size)]),
    "thmm_expm_reversible": (c_status, [c_void_p, c_void_p, c_size, c_void_p, c_size, c_void_p]),
    "thmm_categorical_set_branch_lengths": (c_status, [c_void_p, c_void_p, c_void_p, c_void_p, c_size]),
    "thmm_get_tuning": (c_status, [c_void_p, P(Tuning)]),
    "thmm_set_tuning": (c_status, [c_void_p, P(Tuning)]),
    "thmm_autotune_marginals_device": (
        c_status,
        [c_void_p, c_void_p, c_void_p, c_size, c_void_p, c_void_p, c_void_p, c_void_p, P(Tuning)],
    ),
    "thmm_log_partition_device": (c_status, [c_void_p, c_void_p, c_void_p, c_size, c_void_p, c_void_p]),
    "thmm_static_scratch_bytes": (c_status, [c_void_p, c_size, c_size, c_int, P(c_size)]),
    "thmm_set_path_budget": (c_status, [c_void_p, c_size]),
    "thmm_path_storage": (c_status, [c_void_p, P(c_size), P(c_size)]),
    "thmm_log_partition_device_static": (
        c_status,
        [c_void_p, c_size, c_size, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int],
    ),
    "thmm_marginals_device": (
        c_status,
        [c_void_p, c_void_p, c_void_p, c_size, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "thmm_logz_forward_device": (c_status, [c_void_p, c_void_p, c_void_p, c_size, c_void_p, c_void_p]),
    "thmm_logz_backward_device": (c_status, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "thmm_maximum_device": (c_status, [c_void_p, c_void_p, c_void_p, c_size, c_void_p, c_void_p, c_void_p]),
    "thmm_sample_device": (
        c_status,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_size, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "thmm_synchronize": (c_status, [c_void_p]),
    "thmm_profile_marginals_device": (
        c_status,
        [c_void_p, c_void_p, c_void_p, c_size, c_void_p, c_void_p, c_void_p, c_void_p, P(LaunchTiming), c_size, P(c_size)],
    ),
}

_LIB: ctypes.CDLL | None = None


def library_path() -> Path:
    return Path(os.environ.get("THMM_LIBRARY", LIBRARY_PATH))


def load() -> ctypes.CDLL:
    """Loads libtree_hmm_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: build it with `python -m parallel_tree_hmm_amd.build` "
            "(the HIP library is required; there is no fallback)"
        )
    try:  # share torch's HIP runtime when torch is present
        import torch  # noqa: F401
    except Exception:  # pragma: no cover - torch is optional for the ABI
        pass
    lib = ctypes.CDLL(str(path), mode=ctypes.RTLD_GLOBAL)
    for name, (restype, argtypes) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _LIB = lib
    return lib
