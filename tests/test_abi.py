"""The C-ABI library loads and exports every symbol include/*.h declares;
host-side logic that needs no GPU behaves (CPU)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import parallel_tree_hmm_amd as thmm
from parallel_tree_hmm_amd import _abi

INCLUDE = Path(__file__).resolve().parent.parent / "include"


def declared_functions() -> set[str]:
    names = set()
    for header in INCLUDE.glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names.update(re.findall(r"\b(thmm_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = declared_functions()
    assert len(names) >= 28
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function also has a ctypes prototype
    assert names <= set(_abi.PROTOTYPES), sorted(names - set(_abi.PROTOTYPES))


def test_struct_layouts_match_reference_sizes():
    # src/gpu_backend_impl.inc:39-42
    assert ctypes.sizeof(_abi.Rake) == 16
    assert ctypes.sizeof(_abi.BranchCombination) == 12
    assert ctypes.sizeof(_abi.BranchAbsorption) == 12
    assert ctypes.sizeof(_abi.Compression) == 24


def test_available_returns_a_flag():
    assert thmm.available() in (True, False)


def test_workspace_misuse_is_reported_without_a_gpu():
    ws = thmm.Workspace()
    with pytest.raises(thmm.LogicError):
        ws.inputs(1)  # Inputs before Reserve: logic_error (src/gpu_backend_impl.inc:2084-2085)
    with pytest.raises(thmm.LogicError):
        ws.uniforms(1)
    with pytest.raises(thmm.LogicError):
        ws.categorical_inputs(1)
    plan = thmm.Plan([-1, 0, 0])
    with pytest.raises(thmm.InvalidArgument):
        ws.reserve(plan, 0, 4)  # zero states (src/gpu_backend_impl.inc:1805-1806)
    with pytest.raises(thmm.InvalidArgument):
        ws.reserve(plan, 65, 4)  # above the 64-state kernel limit
    with pytest.raises(thmm.InvalidArgument):
        ws.reserve_categorical(plan, 4, 2, 0, [1])  # categories in [1, 256]
    with pytest.raises(thmm.InvalidArgument):
        ws.reserve_categorical(plan, 4, 2, 3, [2, 1])  # strictly increasing
    # Binary16 edge outputs only for device-resident marginal reservations.
    with pytest.raises(thmm.InvalidArgument):
        ws.reserve_marginals(plan, 4, 2, half_edge_marginals=True)
    with pytest.raises(thmm.InvalidArgument):
        ws.reserve(plan, 4, 2, flags=_abi.THMM_RESERVE_DEVICE_RESIDENT | _abi.THMM_RESERVE_HALF_EDGE_MARGINALS)
    half_marginals = _abi.THMM_RESERVE_MARGINALS | _abi.THMM_RESERVE_DEVICE_RESIDENT | _abi.THMM_RESERVE_HALF_EDGE_MARGINALS
    with pytest.raises(thmm.InvalidArgument):  # categorical reservations drop the device-resident flag
        ws._reserve_categorical(plan, 4, 2, 3, [1], half_marginals, 0, "f32")
    ws.close()


@pytest.mark.skipif(thmm.available(), reason="checks the no-device error path")
def test_reserve_without_device_raises_runtime_error():
    ws = thmm.Workspace()
    with pytest.raises(thmm.AcceleratorRuntimeError):
        ws.reserve(thmm.Plan([-1, 0]), 2, 2)


def test_plan_round_trip_through_arrays():
    plan = thmm.Plan(np.array([-1, 0, 0, 1, 3]))
    copy = thmm.Plan.from_arrays(plan.arrays)
    assert np.array_equal(copy.rakes, plan.rakes)
    assert np.array_equal(copy.compressions, plan.compressions)
    assert np.array_equal(copy.primitive_batches, plan.primitive_batches)
