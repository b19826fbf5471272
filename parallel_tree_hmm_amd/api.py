"""Python mirror of the reference accelerator API (tree_hmm::rocm).

Names, argument meaning and error behaviour follow
include/tree_hmm/accelerator.h and include/tree_hmm/detail/gpu_backend_api.inc
of joaospinto/parallel_tree_hmm; every call goes through the C-ABI of
libtree_hmm_amd.so (include/tree_hmm_amd.h). Result arrays are numpy views of
workspace-owned pinned memory and stay valid until the workspace's next call,
Reserve or destruction, as the reference's spans do.
"""

from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _abi

# ---------------------------------------------------------------------------
# Errors: the C++ hierarchy (invalid_argument, length_error and domain_error
# derive from logic_error).
# ---------------------------------------------------------------------------


class TreeHmmError(Exception):
    """Base of every error raised through the C-ABI."""


class LogicError(TreeHmmError):
    """std::logic_error."""


class InvalidArgument(LogicError, ValueError):
    """std::invalid_argument."""


class LengthError(LogicError):
    """std::length_error."""


class DomainError(LogicError, ArithmeticError):
    """std::domain_error."""


class AcceleratorRuntimeError(TreeHmmError, RuntimeError):
    """std::runtime_error (HIP failures)."""


_ERRORS = {
    _abi.THMM_INVALID_ARGUMENT: InvalidArgument,
    _abi.THMM_LOGIC_ERROR: LogicError,
    _abi.THMM_LENGTH_ERROR: LengthError,
    _abi.THMM_DOMAIN_ERROR: DomainError,
    _abi.THMM_RUNTIME_ERROR: AcceleratorRuntimeError,
}


def _lib():
    return _abi.load()


def _check(status: int) -> None:
    if status != _abi.THMM_OK:
        message = _lib().thmm_last_error().decode(errors="replace")
        raise _ERRORS.get(status, TreeHmmError)(message)


# ---------------------------------------------------------------------------
# Plan
# ---------------------------------------------------------------------------


def _array(ptr, count: int, dtype) -> np.ndarray:
    if count == 0:
        return np.zeros(0, dtype=dtype)
    raw = ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint8))
    nbytes = count * np.dtype(dtype).itemsize
    return np.frombuffer((ctypes.c_uint8 * nbytes).from_address(ctypes.addressof(raw.contents)), dtype=dtype)


_RAKE = np.dtype([("leaf", "<u4"), ("parent", "<u4"), ("edge", "<u4"), ("branch", "<u4")])
_COMBINATION = np.dtype([("destination", "<u4"), ("source", "<u4"), ("stage", "<u4")])
_ABSORPTION = np.dtype([("parent", "<u4"), ("branch", "<u4"), ("round", "<u4")])
_COMPRESSION = np.dtype(
    [("left_edge", "<u4"), ("middle", "<u4"), ("right_edge", "<u4"), ("parent", "<u4"), ("child", "<u4"), ("tape", "<u4")]
)
_BATCH = np.dtype([("primitive", "<u4"), ("offset", "<u4"), ("count", "<u4")])


class Plan:
    """Rake-compress plan of a tree or forest (btrc::MakePlan equivalent).

    ``parents[i] == -1`` marks a root; several roots form a forest whose
    trees are contracted together (one log partition per root).
    """

    def __init__(self, parents: Sequence[int]):
        parents = np.ascontiguousarray(np.asarray(parents, dtype=np.int64))
        handle = ctypes.c_void_p()
        _check(
            _lib().thmm_plan_create(
                parents.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), parents.size, ctypes.byref(handle)
            )
        )
        self._handle = handle
        self._load_arrays()

    @classmethod
    def on_device(cls, parents, device: int = 0) -> "Plan":
        """The same plan, built on HIP device `device` (thmm_plan_create_device).
        `parents` is array-like (host), or a torch int64 tensor on that
        device (read in place)."""
        self = cls.__new__(cls)
        handle = ctypes.c_void_p()
        if hasattr(parents, "is_cuda") and parents.is_cuda:
            import torch

            if parents.dtype != torch.int64 or not parents.is_contiguous():
                raise InvalidArgument("device parents must be a contiguous int64 tensor")
            _check(_lib().thmm_plan_create_device(parents.data_ptr(), parents.numel(), 1, device, ctypes.byref(handle)))
        else:
            host = np.ascontiguousarray(np.asarray(parents, dtype=np.int64))
            _check(_lib().thmm_plan_create_device(host.ctypes.data, host.size, 0, device, ctypes.byref(handle)))
        self._handle = handle
        self._load_arrays()
        return self

    @classmethod
    def from_arrays(cls, arrays: "_abi.PlanArrays") -> "Plan":
        self = cls.__new__(cls)
        handle = ctypes.c_void_p()
        _check(_lib().thmm_plan_create_from_arrays(ctypes.byref(arrays), ctypes.byref(handle)))
        self._handle = handle
        self._load_arrays()
        return self

    def _load_arrays(self) -> None:
        arrays = _abi.PlanArrays()
        _check(_lib().thmm_plan_get_arrays(self._handle, ctypes.byref(arrays)))
        self.arrays = arrays
        self.num_nodes = arrays.num_nodes
        self.num_edges = arrays.num_edges
        self.num_branches = arrays.num_branches
        self.num_compressions = arrays.num_compressions
        self.num_rounds = arrays.num_rounds
        self.roots = _array(arrays.roots, arrays.num_roots, np.uint32)
        self.edge_parents = _array(arrays.edge_parents, arrays.num_edges, np.uint32)
        self.edge_children = _array(arrays.edge_children, arrays.num_edges, np.uint32)
        self.rakes = _array(arrays.rakes, arrays.num_rakes, _RAKE)
        self.branch_combinations = _array(arrays.combinations, arrays.num_combinations, _COMBINATION)
        self.branch_absorptions = _array(arrays.absorptions, arrays.num_absorptions, _ABSORPTION)
        self.compressions = _array(arrays.compressions, arrays.num_compressions, _COMPRESSION)
        self.primitive_batches = _array(arrays.batches, arrays.num_batches, _BATCH)
        self.round_batch_offsets = _array(arrays.round_batch_offsets, arrays.num_rounds + 1, np.uint32)

    @property
    def root(self) -> int:
        if self.roots.size != 1:
            raise InvalidArgument("plan has more than one root")
        return int(self.roots[0])

    @property
    def identity(self) -> int:
        return int(self._handle.value)

    def rounds(self) -> list[dict]:
        """Per round: the batch index ranges of each primitive, including the
        branch_reduction_stages used by tests/accelerator_test.h:241."""
        out = []
        for r in range(self.num_rounds):
            lo, hi = int(self.round_batch_offsets[r]), int(self.round_batch_offsets[r + 1])
            kinds = [int(self.primitive_batches[i]["primitive"]) for i in range(lo, hi)]
            out.append(
                {
                    "batches": list(range(lo, hi)),
                    "branch_reduction_stages": [
                        i for i, k in zip(range(lo, hi), kinds) if k == _abi.THMM_BRANCH_COMBINATION
                    ],
                }
            )
        return out

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value:
            try:
                _lib().thmm_plan_destroy(handle)
            except Exception:
                pass
            self._handle = None


def make_plan(parents: Sequence[int]) -> Plan:
    return Plan(parents)


# ---------------------------------------------------------------------------
# Views (include/tree_hmm/accelerator.h)
# ---------------------------------------------------------------------------


class CategoricalInputUpdate(enum.IntEnum):
    kAll = 0
    kFactors = 1
    kNone = 2


@dataclass
class AcceleratorTimings:
    upload_ms: float = 0.0
    kernel_ms: float = 0.0
    download_ms: float = 0.0
    wall_ms: float = 0.0


@dataclass
class BatchedModelView:
    plan: Plan
    states: int
    batch: int
    node_potentials: np.ndarray  # [batch, node, state]
    edge_potentials: np.ndarray  # [edge, parent, child]


# Workspace.inputs() returns the mutable form; the same dataclass serves both.
MutableBatchedModelView = BatchedModelView


@dataclass
class BatchedCategoricalModelView:
    plan: Plan
    states: int
    batch: int
    categories: int
    observation_nodes: np.ndarray  # strictly increasing node indices
    observations: np.ndarray  # [batch, observation node] uint8
    root_potential: np.ndarray  # [state]
    emission_potentials: np.ndarray  # [category, state]
    edge_potentials: np.ndarray  # [edge, parent, child]


MutableBatchedCategoricalModelView = BatchedCategoricalModelView


@dataclass
class PartitionView:
    values: np.ndarray
    timings: AcceleratorTimings


@dataclass
class BatchedMaximumAssignmentView:
    log_weights: np.ndarray
    states: np.ndarray
    timings: AcceleratorTimings


@dataclass
class BatchedPosteriorSampleView:
    states: np.ndarray
    timings: AcceleratorTimings


@dataclass
class BatchedMarginalView:
    log_partitions: np.ndarray
    nodes: np.ndarray
    edges: np.ndarray
    timings: AcceleratorTimings


def available() -> bool:
    """tree_hmm::rocm::Available()."""
    return bool(_lib().thmm_available())


def device_description(device: int = 0) -> str:
    buffer = ctypes.create_string_buffer(256)
    _check(_lib().thmm_device_description(device, buffer, len(buffer)))
    return buffer.value.decode()


_PRECISIONS = {"f32": (_abi.THMM_F32, np.float32), "f64": (_abi.THMM_F64, np.float64)}

_HIP_NULL_STREAM = ctypes.c_void_p(2**64 - 1)  # THMM_HIP_NULL_STREAM


def static_scratch_bytes(plan: Plan, states: int, batch: int, precision: str = "f32") -> int:
    """Bytes of contraction state log_partition_device_static needs for this
    shape (at most _abi.THMM_STATIC_SCRATCH_BYTES are available)."""
    out = ctypes.c_size_t()
    _check(_lib().thmm_static_scratch_bytes(plan._handle, states, batch, _PRECISIONS[precision][0], ctypes.byref(out)))
    return out.value


def log_partition_device_static(plan: Plan, states: int, batch: int, nodes_ptr: int, edges_ptr: int, log_z_ptr: int,
                                precision: str = "f32", stream=0, device: int = 0) -> None:
    """thmm_log_partition_device_static: log Z of device-resident potentials
    with no workspace and no device allocation (state in the library's static
    per-GPU array; calls that overlap on one GPU race). stream 0 or None is
    HIP's null stream."""
    _check(
        _lib().thmm_log_partition_device_static(
            plan._handle, states, batch, _PRECISIONS[precision][0], nodes_ptr, edges_ptr, log_z_ptr, stream or None,
            device
        )
    )


def _stream(stream):
    """Maps a hipStream_t handle to the C-ABI convention (NULL means the
    workspace's own stream, so HIP's null stream 0 gets its sentinel)."""
    if stream is None:
        return None
    return _HIP_NULL_STREAM if int(stream) == 0 else ctypes.c_void_p(int(stream))


class Workspace:
    """tree_hmm::rocm::Workspace: device + pinned storage for prepared calls."""

    def __init__(self):
        handle = ctypes.c_void_p()
        _check(_lib().thmm_workspace_create(ctypes.byref(handle)))
        self._handle = handle
        self._plan: Optional[Plan] = None
        self.dtype = np.float32
        self.states = 0
        self.capacity = 0
        self.categories = 0
        self.observation_nodes = np.zeros(0, np.uint32)

    def close(self) -> None:
        if self._handle is not None and self._handle.value:
            _lib().thmm_workspace_destroy(self._handle)
        self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- Reserve family -------------------------------------------------
    def _reserve(self, plan: Plan, states: int, batch_capacity: int, flags: int, device: int, precision: str):
        code, dtype = _PRECISIONS[precision]
        self._plan = None
        _check(
            _lib().thmm_reserve(
                self._handle, plan._handle, ctypes.c_void_p(plan.identity), states, batch_capacity, code, flags, device
            )
        )
        self._plan, self.dtype, self.states, self.capacity = plan, dtype, states, batch_capacity

    def _reserve_categorical(self, plan, states, batch_capacity, categories, observation_nodes, flags, device, precision):
        code, dtype = _PRECISIONS[precision]
        obs = np.ascontiguousarray(np.asarray(observation_nodes, dtype=np.uint32))
        self._plan = None
        _check(
            _lib().thmm_reserve_categorical(
                self._handle,
                plan._handle,
                ctypes.c_void_p(plan.identity),
                states,
                batch_capacity,
                categories,
                obs.ctypes.data if obs.size else None,
                obs.size,
                code,
                flags,
                device,
            )
        )
        self._plan, self.dtype, self.states, self.capacity = plan, dtype, states, batch_capacity
        self.categories, self.observation_nodes = categories, obs

    def set_path_budget(self, nbytes: int) -> None:
        """Path-storage budget (bytes) for the next reservations; 0 = none.
        Forests above it run in tree groups (see thmm_set_path_budget)."""
        _check(_lib().thmm_set_path_budget(self._handle, int(nbytes)))

    def path_storage(self) -> tuple:
        """(tree groups of the reservation, 0 if not grouped; bytes of path
        storage held)."""
        groups, nbytes = ctypes.c_size_t(), ctypes.c_size_t()
        _check(_lib().thmm_path_storage(self._handle, ctypes.byref(groups), ctypes.byref(nbytes)))
        return groups.value, nbytes.value

    def reserve(self, plan, states, batch_capacity, device=0, precision="f32", flags=0):
        self._reserve(plan, states, batch_capacity, flags, device, precision)

    def reserve_maximum(self, plan, states, batch_capacity, device=0, precision="f32"):
        self._reserve(plan, states, batch_capacity, _abi.THMM_RESERVE_MAXIMUM, device, precision)

    def reserve_sampling(self, plan, states, batch_capacity, device=0, precision="f32"):
        self._reserve(plan, states, batch_capacity, _abi.THMM_RESERVE_SAMPLING, device, precision)

    def reserve_marginals(self, plan, states, batch_capacity, device=0, precision="f32", device_resident=False,
                          graphs=False, half_edge_marginals=False):
        """ReserveMarginals. device_resident: no pinned host mirrors (device
        entry points only); graphs: device entry points replay a captured
        hipGraph of their launch sequence; half_edge_marginals (with
        device_resident): the device entry points write edge marginals and
        edge gradients as IEEE binary16."""
        flags = _abi.THMM_RESERVE_MARGINALS | (_abi.THMM_RESERVE_DEVICE_RESIDENT if device_resident else 0)
        flags |= _abi.THMM_RESERVE_GRAPH if graphs else 0
        flags |= _abi.THMM_RESERVE_HALF_EDGE_MARGINALS if half_edge_marginals else 0
        self._reserve(plan, states, batch_capacity, flags, device, precision)

    def reserve_categorical(self, plan, states, batch_capacity, categories, observation_nodes, device=0, precision="f32"):
        self._reserve_categorical(plan, states, batch_capacity, categories, observation_nodes, 0, device, precision)

    def reserve_categorical_maximum(self, plan, states, batch_capacity, categories, observation_nodes, device=0, precision="f32"):
        self._reserve_categorical(
            plan, states, batch_capacity, categories, observation_nodes, _abi.THMM_RESERVE_MAXIMUM, device, precision
        )

    def reserve_categorical_sampling(self, plan, states, batch_capacity, categories, observation_nodes, device=0, precision="f32"):
        self._reserve_categorical(
            plan, states, batch_capacity, categories, observation_nodes, _abi.THMM_RESERVE_SAMPLING, device, precision
        )

    def reserve_categorical_marginals(self, plan, states, batch_capacity, categories, observation_nodes, device=0, precision="f32"):
        self._reserve_categorical(
            plan, states, batch_capacity, categories, observation_nodes, _abi.THMM_RESERVE_MARGINALS, device, precision
        )

    # -- pinned inputs ---------------------------------------------------
    def inputs(self, batch: Optional[int] = None) -> BatchedModelView:
        batch = self.capacity if batch is None else batch
        nodes, node_count = ctypes.c_void_p(), ctypes.c_size_t()
        edges, edge_count = ctypes.c_void_p(), ctypes.c_size_t()
        _check(
            _lib().thmm_inputs(
                self._handle, batch, ctypes.byref(nodes), ctypes.byref(node_count), ctypes.byref(edges), ctypes.byref(edge_count)
            )
        )
        plan = self._plan
        return BatchedModelView(
            plan,
            self.states,
            batch,
            _array(nodes, node_count.value, self.dtype).reshape(batch, plan.num_nodes, self.states),
            _array(edges, edge_count.value, self.dtype).reshape(plan.num_edges, self.states, self.states),
        )

    def categorical_inputs(self, batch: Optional[int] = None) -> BatchedCategoricalModelView:
        batch = self.capacity if batch is None else batch
        obs, obs_count = ctypes.c_void_p(), ctypes.c_size_t()
        root, emissions, emission_count = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_size_t()
        edges, edge_count = ctypes.c_void_p(), ctypes.c_size_t()
        _check(
            _lib().thmm_categorical_inputs(
                self._handle,
                batch,
                ctypes.byref(obs),
                ctypes.byref(obs_count),
                ctypes.byref(root),
                ctypes.byref(emissions),
                ctypes.byref(emission_count),
                ctypes.byref(edges),
                ctypes.byref(edge_count),
            )
        )
        plan = self._plan
        n_obs = self.observation_nodes.size
        return BatchedCategoricalModelView(
            plan,
            self.states,
            batch,
            self.categories,
            self.observation_nodes,
            _array(obs, obs_count.value, np.uint8).reshape(batch, n_obs),
            _array(root, self.states, self.dtype),
            _array(emissions, emission_count.value, self.dtype).reshape(self.categories, self.states),
            _array(edges, edge_count.value, self.dtype).reshape(plan.num_edges, self.states, self.states),
        )

    def uniforms(self, batch: Optional[int] = None) -> np.ndarray:
        batch = self.capacity if batch is None else batch
        ptr, count = ctypes.c_void_p(), ctypes.c_size_t()
        _check(_lib().thmm_uniforms(self._handle, batch, ctypes.byref(ptr), ctypes.byref(count)))
        return _array(ptr, count.value, self.dtype).reshape(batch, self._plan.num_nodes)

    # -- device-resident entry points --------------------------------------
    # `stream` is a hipStream_t handle (torch: stream.cuda_stream); 0 is
    # HIP's null stream. None runs on the workspace's own stream.
    def log_partition_device(self, nodes_ptr: int, edges_ptr: int, batch: int, log_z_ptr: int, stream=None):
        _check(_lib().thmm_log_partition_device(self._handle, nodes_ptr, edges_ptr, batch, log_z_ptr, _stream(stream)))

    def marginals_device(self, nodes_ptr, edges_ptr, batch, log_z_ptr, node_out_ptr, edge_out_ptr, stream=None):
        _check(
            _lib().thmm_marginals_device(
                self._handle, nodes_ptr, edges_ptr, batch, log_z_ptr, node_out_ptr, edge_out_ptr, _stream(stream)
            )
        )

    def logz_forward_device(self, log_nodes_ptr, log_edges_ptr, batch, log_z_ptr, stream=None):
        _check(
            _lib().thmm_logz_forward_device(self._handle, log_nodes_ptr, log_edges_ptr, batch, log_z_ptr, _stream(stream))
        )

    def logz_backward_device(self, grad_ptr, grad_nodes_ptr, grad_edges_ptr, reduce_edges: bool, stream=None):
        _check(
            _lib().thmm_logz_backward_device(
                self._handle, grad_ptr or None, grad_nodes_ptr, grad_edges_ptr, int(reduce_edges), _stream(stream)
            )
        )

    def maximum_device(self, nodes_ptr, edges_ptr, batch, log_weights_ptr, states_ptr, stream=None):
        """thmm_maximum_device: MAP log weights (double) and states (uint32
        [batch, node]) into device buffers."""
        _check(
            _lib().thmm_maximum_device(
                self._handle, nodes_ptr, edges_ptr, batch, log_weights_ptr, states_ptr, _stream(stream)
            )
        )

    def sample_device(self, nodes_ptr, edges_ptr, uniforms_ptr, batch, log_z_ptr, states_ptr, failure_ptr,
                      stream=None):
        """thmm_sample_device: posterior states (uint32 [batch, node]) from
        device variates; log Z (double) and a uint32 failure flag."""
        _check(
            _lib().thmm_sample_device(
                self._handle, nodes_ptr, edges_ptr, uniforms_ptr, batch, log_z_ptr, states_ptr, failure_ptr,
                _stream(stream)
            )
        )

    def profile_marginals_device(self, nodes_ptr, edges_ptr, batch, log_z_ptr, node_out_ptr, edge_out_ptr, stream=None):
        """thmm_marginals_device with per-launch HIP event timing; returns a
        list of (phase, primitive, count, batch, ms) tuples."""
        capacity = 4 * (self._plan.primitive_batches.size + 8)
        timings = (_abi.LaunchTiming * capacity)()
        count = ctypes.c_size_t()
        _check(
            _lib().thmm_profile_marginals_device(
                self._handle, nodes_ptr, edges_ptr, batch, log_z_ptr, node_out_ptr, edge_out_ptr, _stream(stream),
                timings, capacity, ctypes.byref(count)
            )
        )
        kept = timings[: min(count.value, capacity)]
        # work units per launch (compression groups for grouped levels)
        self.last_profile_ops = [t.ops for t in kept]
        return [(t.phase, t.primitive, t.count, t.batch, t.ms) for t in kept]

    # -- launch tuning -----------------------------------------------------
    _TUNING_FIELDS = ("compress_items_forward", "compress_items_reverse", "rake_block", "gather_chunk",
                      "compress_depth", "rake_log_edges", "persistent", "input_leaves")

    def tuning(self) -> dict:
        t = _abi.Tuning()
        _check(_lib().thmm_get_tuning(self._handle, ctypes.byref(t)))
        return {f: getattr(t, f) for f in self._TUNING_FIELDS}

    def set_tuning(self, **values) -> None:
        current = self.tuning()
        unknown = set(values) - set(current)
        if unknown:
            raise InvalidArgument(f"unknown tuning fields {sorted(unknown)}")
        current.update(values)
        t = _abi.Tuning(*(int(current[f]) for f in self._TUNING_FIELDS), 0.0)
        _check(_lib().thmm_set_tuning(self._handle, ctypes.byref(t)))

    def autotune_marginals_device(self, nodes_ptr, edges_ptr, batch, log_z_ptr, node_out_ptr, edge_out_ptr,
                                  stream=None) -> dict:
        """Times the launch variants on these inputs, keeps the fastest and
        returns it with its milliseconds per marginal pass ("ms")."""
        t = _abi.Tuning()
        _check(
            _lib().thmm_autotune_marginals_device(
                self._handle, nodes_ptr, edges_ptr, batch, log_z_ptr, node_out_ptr, edge_out_ptr, _stream(stream),
                ctypes.byref(t)
            )
        )
        out = {f: getattr(t, f) for f in self._TUNING_FIELDS}
        out["ms"] = t.ms
        return out

    def synchronize(self) -> None:
        _check(_lib().thmm_synchronize(self._handle))

    # -- branch-length gradient pair (categorical) --------------------------
    def categorical_logz_forward(self, model: "BatchedCategoricalModelView", update=None) -> np.ndarray:
        """Forward half: FP64 log partitions [batch * roots] of a categorical
        model (edge factors P_e = expm(Q t_e)); keeps the contraction state
        for branch_length_backward."""
        m, keep = _categorical(model, self)
        update = CategoricalInputUpdate.kAll if update is None else update
        ptr, count = ctypes.c_void_p(), ctypes.c_size_t()
        _check(_lib().thmm_categorical_logz_forward(self._handle, ctypes.byref(m), int(update), ctypes.byref(ptr), ctypes.byref(count)))
        del keep
        return _array(ptr, count.value, np.float64).copy()

    def categorical_log_likelihood(self, model: "BatchedCategoricalModelView", update=None) -> np.ndarray:
        """Forward half with the sites summed on the device: FP64 [roots],
        sum over the batch of log Z; keeps the contraction state for
        branch_length_backward (grad_log_z None then differentiates the
        sum)."""
        m, keep = _categorical(model, self)
        update = CategoricalInputUpdate.kAll if update is None else update
        ptr, count = ctypes.c_void_p(), ctypes.c_size_t()
        _check(_lib().thmm_categorical_log_likelihood(self._handle, ctypes.byref(m), int(update), ctypes.byref(ptr),
                                                      ctypes.byref(count)))
        del keep
        return _array(ptr, count.value, np.float64).copy()

    def branch_length_backward(self, rate_matrix, grad_log_z=None) -> np.ndarray:
        """Backward half: d (sum_b g_b log Z_b) / d t_e for every edge, FP64
        [num_edges], given the K x K rate matrix Q and the seeds g (None =
        all ones)."""
        k = self.states
        q = np.ascontiguousarray(np.asarray(rate_matrix, dtype=np.float64))
        if q.shape != (k, k):
            raise InvalidArgument(f"rate matrix must be {k} x {k}")
        g = None
        if grad_log_z is not None:
            g = np.ascontiguousarray(np.asarray(grad_log_z, dtype=np.float64)).ravel()
        ptr, count = ctypes.c_void_p(), ctypes.c_size_t()
        _check(
            _lib().thmm_branch_length_backward(
                self._handle, None if g is None else g.ctypes.data, q.ctypes.data, ctypes.byref(ptr), ctypes.byref(count)
            )
        )
        return _array(ptr, count.value, np.float64).copy()

    def categorical_set_branch_lengths(self, rate_matrix, stationary, branch_lengths) -> None:
        """Device edge factors P_e = expm(Q t_e) for a rate matrix Q
        reversible with respect to `stationary`, one branch length per plan
        edge; follow with calls using CategoricalInputUpdate.kNone."""
        k = self.states
        q = np.ascontiguousarray(np.asarray(rate_matrix, dtype=np.float64))
        pi = np.ascontiguousarray(np.asarray(stationary, dtype=np.float64)).ravel()
        t = np.ascontiguousarray(np.asarray(branch_lengths, dtype=np.float64)).ravel()
        if q.shape != (k, k) or pi.size != k:
            raise InvalidArgument(f"rate matrix must be {k} x {k} and the stationary law of length {k}")
        _check(_lib().thmm_categorical_set_branch_lengths(self._handle, q.ctypes.data, pi.ctypes.data, t.ctypes.data,
                                                          t.size))

    def log_partitions_f64(self) -> np.ndarray:
        ptr, count = ctypes.c_void_p(), ctypes.c_size_t()
        _check(_lib().thmm_result_log_partitions_f64(self._handle, ctypes.byref(ptr), ctypes.byref(count)))
        return _array(ptr, count.value, np.float64)


# ---------------------------------------------------------------------------
# Prepared calls (free functions, as in gpu_backend_api.inc:97-142)
# ---------------------------------------------------------------------------


def expm_reversible(rate_matrix, stationary, branch_lengths) -> np.ndarray:
    """P_e = expm(Q t_e), FP64 [edges, K, K], for a rate matrix Q reversible
    with respect to `stationary` (host computation, thmm_expm_reversible)."""
    q = np.ascontiguousarray(np.asarray(rate_matrix, dtype=np.float64))
    pi = np.ascontiguousarray(np.asarray(stationary, dtype=np.float64)).ravel()
    t = np.ascontiguousarray(np.asarray(branch_lengths, dtype=np.float64)).ravel()
    k = pi.size
    if q.shape != (k, k):
        raise InvalidArgument(f"rate matrix must be {k} x {k}")
    out = np.empty((t.size, k, k), np.float64)
    _check(_lib().thmm_expm_reversible(q.ctypes.data, pi.ctypes.data, k, t.ctypes.data, t.size, out.ctypes.data))
    return out


def _buffer(array, dtype, name: str):
    """Returns (pointer, count, keepalive) for a numpy input, converting to
    the workspace dtype when needed (a converted copy is staged)."""
    a = np.asarray(array)
    if a.dtype != dtype or not a.flags.c_contiguous:
        a = np.ascontiguousarray(a, dtype=dtype)
    return (a.ctypes.data if a.size else None), a.size, a


def _dense(model: BatchedModelView, ws: Workspace):
    nodes, node_count, keep_n = _buffer(model.node_potentials, ws.dtype, "node")
    edges, edge_count, keep_e = _buffer(model.edge_potentials, ws.dtype, "edge")
    m = _abi.DenseModel(
        ctypes.c_void_p(model.plan.identity), model.states, model.batch, nodes, node_count, edges, edge_count
    )
    return m, (keep_n, keep_e)


def _categorical(model: BatchedCategoricalModelView, ws: Workspace):
    obs_nodes = np.ascontiguousarray(np.asarray(model.observation_nodes, dtype=np.uint32))
    obs = np.asarray(model.observations)
    if obs.dtype != np.uint8 or not obs.flags.c_contiguous:
        obs = np.ascontiguousarray(obs, dtype=np.uint8)
    root, root_count, keep_r = _buffer(model.root_potential, ws.dtype, "root")
    em, em_count, keep_m = _buffer(model.emission_potentials, ws.dtype, "emission")
    edges, edge_count, keep_e = _buffer(model.edge_potentials, ws.dtype, "edge")
    m = _abi.CategoricalModel(
        ctypes.c_void_p(model.plan.identity),
        model.states,
        model.batch,
        model.categories,
        obs_nodes.ctypes.data if obs_nodes.size else None,
        obs_nodes.size,
        obs.ctypes.data if obs.size else None,
        obs.size,
        root,
        root_count,
        em,
        em_count,
        edges,
        edge_count,
    )
    return m, (obs_nodes, obs, keep_r, keep_m, keep_e)


def _timings(t: _abi.Timings) -> AcceleratorTimings:
    return AcceleratorTimings(t.upload_ms, t.kernel_ms, t.download_ms, t.wall_ms)


def _is_categorical(model) -> bool:
    return isinstance(model, BatchedCategoricalModelView)


def _partition(model, ws: Workspace, log_domain: bool, update) -> PartitionView:
    values, count, t = ctypes.c_void_p(), ctypes.c_size_t(), _abi.Timings()
    if _is_categorical(model):
        m, keep = _categorical(model, ws)
        update = CategoricalInputUpdate.kAll if update is None else update
        _check(
            _lib().thmm_partition_categorical(
                ws._handle, ctypes.byref(m), int(update), int(log_domain), ctypes.byref(values), ctypes.byref(count), ctypes.byref(t)
            )
        )
    else:
        m, keep = _dense(model, ws)
        _check(_lib().thmm_partition(ws._handle, ctypes.byref(m), int(log_domain), ctypes.byref(values), ctypes.byref(count), ctypes.byref(t)))
    del keep
    return PartitionView(_array(values, count.value, ws.dtype), _timings(t))


def partition_function_prepared(model, workspace: Workspace, update=None) -> PartitionView:
    """PartitionFunctionPrepared (dense or categorical)."""
    return _partition(model, workspace, False, update)


def log_partition_function_prepared(model, workspace: Workspace, update=None) -> PartitionView:
    """LogPartitionFunctionPrepared (dense or categorical)."""
    return _partition(model, workspace, True, update)


def maximum_a_posteriori_prepared(model, workspace: Workspace, update=None) -> BatchedMaximumAssignmentView:
    ws = workspace
    weights, states, t = ctypes.c_void_p(), ctypes.c_void_p(), _abi.Timings()
    if _is_categorical(model):
        m, keep = _categorical(model, ws)
        update = CategoricalInputUpdate.kAll if update is None else update
        _check(_lib().thmm_maximum_categorical(ws._handle, ctypes.byref(m), int(update), ctypes.byref(weights), ctypes.byref(states), ctypes.byref(t)))
    else:
        m, keep = _dense(model, ws)
        _check(_lib().thmm_maximum(ws._handle, ctypes.byref(m), ctypes.byref(weights), ctypes.byref(states), ctypes.byref(t)))
    del keep
    roots = model.plan.roots.size
    return BatchedMaximumAssignmentView(
        _array(weights, model.batch * roots, ws.dtype),
        _array(states, model.batch * model.plan.num_nodes, np.uint32),
        _timings(t),
    )


def posterior_sample_prepared(model, uniforms, workspace: Workspace, update=None) -> BatchedPosteriorSampleView:
    ws = workspace
    u, u_count, keep_u = _buffer(uniforms, ws.dtype, "uniform")
    states, t = ctypes.c_void_p(), _abi.Timings()
    if _is_categorical(model):
        m, keep = _categorical(model, ws)
        update = CategoricalInputUpdate.kAll if update is None else update
        _check(_lib().thmm_sample_categorical(ws._handle, ctypes.byref(m), u, u_count, int(update), ctypes.byref(states), ctypes.byref(t)))
    else:
        m, keep = _dense(model, ws)
        _check(_lib().thmm_sample(ws._handle, ctypes.byref(m), u, u_count, ctypes.byref(states), ctypes.byref(t)))
    del keep, keep_u
    return BatchedPosteriorSampleView(_array(states, model.batch * model.plan.num_nodes, np.uint32), _timings(t))


def posterior_marginals_prepared(model, workspace: Workspace, update=None) -> BatchedMarginalView:
    ws = workspace
    logz, nodes, edges, t = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), _abi.Timings()
    if _is_categorical(model):
        m, keep = _categorical(model, ws)
        update = CategoricalInputUpdate.kAll if update is None else update
        _check(
            _lib().thmm_marginals_categorical(
                ws._handle, ctypes.byref(m), int(update), ctypes.byref(logz), ctypes.byref(nodes), ctypes.byref(edges), ctypes.byref(t)
            )
        )
    else:
        m, keep = _dense(model, ws)
        _check(_lib().thmm_marginals(ws._handle, ctypes.byref(m), ctypes.byref(logz), ctypes.byref(nodes), ctypes.byref(edges), ctypes.byref(t)))
    del keep
    plan = model.plan
    k = model.states
    return BatchedMarginalView(
        _array(logz, model.batch * plan.roots.size, ws.dtype),
        _array(nodes, model.batch * plan.num_nodes * k, ws.dtype),
        _array(edges, model.batch * plan.num_edges * k * k, ws.dtype),
        _timings(t),
    )
