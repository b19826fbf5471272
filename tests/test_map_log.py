"""LogOfFloat (parallel_tree_hmm_amd/csrc/log_f32.h): the FP64 log that the
MAP kernels take of FP32 potentials (src/inference.cc:95-98 takes the log of
every potential before the max-product pass).

Checked on the host, from the header itself (g++ -ffp-contract=off: the
operation sequence the device compiles to):
* against the x87 extended-precision logl over a strided sweep of every
  positive finite float: at most 0.504 ulp, and results beyond 0.5 + 2^-10
  ulp rarer than the C library log's own;
* against the C library log (the FP64 oracle's): the same bits for all but a
  few in 10^4 values (the exhaustive sweep, stride 1, differs on 65,685 of
  2,139,095,039 floats, where the library log is itself off by more than
  0.5 ulp 10,666 times and LogOfFloat 28 times);
* exact values the max-product ties rely on: log 1 = 0, log 2^k = -log 2^-k,
  and agreement with mpmath near 1, near powers of two and at the ends of the
  float range.
"""

from __future__ import annotations

import ctypes
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "parallel_tree_hmm_amd" / "csrc"
CPP = ROOT / "tests" / "cpp"
FLAGS = ["-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas", f"-I{CSRC}"]


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    out = tmp_path_factory.mktemp("logf32")
    check = out / "log_check"
    lib = out / "liblogvalues.so"
    subprocess.run(["g++", *FLAGS, str(CPP / "log_check.cc"), "-o", str(check), "-lpthread"], check=True)
    subprocess.run(["g++", *FLAGS, "-shared", "-fPIC", str(CPP / "log_values.cc"), "-o", str(lib)], check=True)
    return check, lib


def host_log(lib, x: np.ndarray) -> np.ndarray:
    dll = ctypes.CDLL(str(lib))
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(x.size, np.float64)
    dll.log_of_floats(x.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(x.size))
    return out.reshape(x.shape)


def test_strided_sweep_against_logl_and_libm(built):
    check, _ = built
    result = subprocess.run([str(check), "997", "3", "8"], capture_output=True, text=True, check=True, timeout=300)
    values, differ, worst, ours_far, libm_far = result.stdout.split()
    values, differ, ours_far, libm_far = int(values), int(differ), int(ours_far), int(libm_far)
    print(f"{values} floats: {differ} differ from log, worst {worst} ulp, "
          f"beyond 0.5+2^-10 ulp: ours {ours_far}, log's {libm_far}")
    assert values > 2_000_000
    assert float(worst) <= 0.504
    assert ours_far <= libm_far
    assert differ <= 1e-4 * values


def test_exact_values(built):
    _, lib = built
    k = np.arange(-149, 128)
    powers = np.ldexp(np.float32(1.0), k).astype(np.float32)
    logs = host_log(lib, powers)
    assert logs[k == 0][0] == 0.0
    for i, kk in enumerate(k):
        if -kk in set(k.tolist()):
            assert logs[i] == -logs[k == -kk][0]
    import mpmath

    mpmath.mp.prec = 120
    probes = [1.0, np.nextafter(np.float32(1), np.float32(2)), np.nextafter(np.float32(1), np.float32(0)),
              np.float32(0.75), np.float32(1.5), np.nextafter(np.float32(2), np.float32(0)),
              np.float32(np.finfo(np.float32).tiny), np.float32(1e-45), np.float32(np.finfo(np.float32).max),
              np.float32(math.e), np.float32(10.0), np.float32(0.1)]
    probes = np.array(probes, np.float32)
    got = host_log(lib, probes)
    for x, y in zip(probes, got):
        exact = mpmath.log(mpmath.mpf(float(x)))
        assert float(exact) == y, (x, y, exact)
    assert host_log(lib, np.array([0.0, -1.0, np.nan], np.float32)).tolist() == [-math.inf] * 3
    assert host_log(lib, np.array([np.inf], np.float32))[0] == math.inf


def test_monotone_on_random_floats(built):
    _, lib = built
    rng = np.random.default_rng(3)
    bits = np.sort(rng.integers(1, 0x7F800000, size=1_000_000, dtype=np.uint32))
    logs = host_log(lib, bits.view(np.float32))
    assert np.all(np.diff(logs) >= 0)
