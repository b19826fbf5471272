"""Runs the C++ drop-in conformance programs (tests/cpp/conformance.cc,
built by parallel_tree_hmm_amd/build.py for Scalar = double and float)."""

from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

HERE = Path(__file__).resolve().parent / "cpp"


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_cpp_layer_conformance(require_gpu, precision):
    binary = HERE / f"conformance_{precision}"
    assert binary.exists(), f"{binary} is not built (python -m parallel_tree_hmm_amd.build)"
    result = subprocess.run([str(binary)], capture_output=True, text=True, timeout=100)
    print(result.stdout, result.stderr)
    assert result.returncode == 0, result.stderr
    assert "conformance passed" in result.stdout
