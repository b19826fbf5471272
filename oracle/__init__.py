"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg as the checker; the product never loads it.
Each function restates joaospinto/parallel_tree_hmm src/inference.cc (see
oracle.h for line citations) over a plan's flat arrays.

Pinning (tests/test_oracle.py): brute-force enumeration of Z, marginals and
the MAP assignment on random tiny trees and forests, the literal fixtures
of tests/inference_test.cc:159-193, the closed forms of
tests/inference_test.cc:286-341, central finite differences of log Z, and an
exhaustive conditional sampler that follows the plan's expansion order.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "liboracle.so"

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not LIBRARY.exists():
            raise RuntimeError(f"{LIBRARY} is missing: run `make -C oracle`")
        lib = ctypes.CDLL(str(LIBRARY))
        d = ctypes.POINTER(ctypes.c_double)
        u = ctypes.POINTER(ctypes.c_uint32)
        v = ctypes.c_void_p
        s = ctypes.c_size_t
        protos = {
            "oracle_partition": [v, s, d, d, d],
            "oracle_log_partition": [v, s, d, d, d],
            "oracle_marginals": [v, s, d, d, d, d, d],
            "oracle_map": [v, s, d, d, d, u],
            "oracle_sample": [v, s, d, d, d, u, d],
            "oracle_brute_force": [v, s, d, d, d, d, d, d, u],
            "oracle_brute_conditional_sample": [v, s, d, d, d, u],
            "oracle_marginals_batched": [v, s, s, d, d, d, d, d, s],
        }
        for name, args in protos.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        lib.oracle_assignment_weight.argtypes = [v, s, d, d, u]
        lib.oracle_assignment_weight.restype = ctypes.c_double
        _LIB = lib
    return _LIB


class OracleError(RuntimeError):
    pass


class OracleDomainError(OracleError):
    pass


def _check(rc: int) -> None:
    if rc == 4:
        raise OracleDomainError("domain error")
    if rc != 0:
        raise OracleError(f"oracle failed with code {rc}")


def _d(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _arrays(plan):
    return ctypes.cast(ctypes.pointer(plan.arrays), ctypes.c_void_p)


def partition(plan, k, nodes, edges) -> np.ndarray:
    out = np.zeros(plan.roots.size)
    n, e = _f64(nodes), _f64(edges)
    _check(_lib().oracle_partition(_arrays(plan), k, _d(n), _d(e), _d(out)))
    return out


def log_partition(plan, k, nodes, edges) -> np.ndarray:
    out = np.zeros(plan.roots.size)
    n, e = _f64(nodes), _f64(edges)
    _check(_lib().oracle_log_partition(_arrays(plan), k, _d(n), _d(e), _d(out)))
    return out


def marginals(plan, k, nodes, edges):
    """Returns (log_z [roots], nodes [N, K], edges [E, K, K])."""
    n, e = _f64(nodes), _f64(edges)
    log_z = np.zeros(plan.roots.size)
    nm = np.zeros((plan.num_nodes, k))
    em = np.zeros((plan.num_edges, k, k))
    _check(_lib().oracle_marginals(_arrays(plan), k, _d(n), _d(e), _d(log_z), _d(nm), _d(em)))
    return log_z, nm, em


def marginals_batched(plan, k, nodes, edges, threads=1):
    """Batch items [B, N, K] with shared edges, on `threads` host threads."""
    n, e = _f64(nodes), _f64(edges)
    b = n.shape[0]
    log_z = np.zeros((b, plan.roots.size))
    nm = np.zeros((b, plan.num_nodes, k))
    em = np.zeros((b, plan.num_edges, k, k))
    _check(_lib().oracle_marginals_batched(_arrays(plan), k, b, _d(n), _d(e), _d(log_z), _d(nm), _d(em), threads))
    return log_z, nm, em


def maximum(plan, k, nodes, edges):
    """Returns (log_weight [roots], states [N])."""
    n, e = _f64(nodes), _f64(edges)
    lw = np.zeros(plan.roots.size)
    states = np.zeros(plan.num_nodes, dtype=np.uint32)
    _check(_lib().oracle_map(_arrays(plan), k, _d(n), _d(e), _d(lw), _u(states)))
    return lw, states


def sample(plan, k, nodes, edges, uniforms):
    """Returns (states [N], boundary distance [N])."""
    n, e, u = _f64(nodes), _f64(edges), _f64(uniforms)
    states = np.zeros(plan.num_nodes, dtype=np.uint32)
    dist = np.zeros(plan.num_nodes)
    _check(_lib().oracle_sample(_arrays(plan), k, _d(n), _d(e), _d(u), _u(states), _d(dist)))
    return states, dist


def brute_force(plan, k, nodes, edges):
    """Returns (Z, nodes [N,K], edges [E,K,K], max weight, argmax states)."""
    n, e = _f64(nodes), _f64(edges)
    z, top = np.zeros(1), np.zeros(1)
    nm = np.zeros((plan.num_nodes, k))
    em = np.zeros((plan.num_edges, k, k))
    best = np.zeros(plan.num_nodes, dtype=np.uint32)
    _check(_lib().oracle_brute_force(_arrays(plan), k, _d(n), _d(e), _d(z), _d(nm), _d(em), _d(top), _u(best)))
    return z[0], nm, em, top[0], best


def brute_conditional_sample(plan, k, nodes, edges, uniforms) -> np.ndarray:
    n, e, u = _f64(nodes), _f64(edges), _f64(uniforms)
    states = np.zeros(plan.num_nodes, dtype=np.uint32)
    _check(_lib().oracle_brute_conditional_sample(_arrays(plan), k, _d(n), _d(e), _d(u), _u(states)))
    return states


def assignment_weight(plan, k, nodes, edges, states) -> float:
    n, e = _f64(nodes), _f64(edges)
    s = np.ascontiguousarray(np.asarray(states, dtype=np.uint32))
    return _lib().oracle_assignment_weight(_arrays(plan), k, _d(n), _d(e), _u(s))
