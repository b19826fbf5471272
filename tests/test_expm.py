"""Host expm of reversible rate matrices (thmm_expm_reversible), the
factors the device writes as categorical edge factors: checked against
scipy.linalg.expm (a Pade method, independent of the eigendecomposition),
the Jukes-Cantor closed form, stochastic rows and P(0) = I, and the
argument checks."""

from __future__ import annotations

import numpy as np
import pytest
from scipy.linalg import expm

import parallel_tree_hmm_amd as thmm
from conftest import reversible_rate_matrix, stationary_of


@pytest.mark.parametrize("k", [1, 2, 4, 20, 61])
def test_matches_scipy_expm(k):
    rng = np.random.default_rng(k)
    q = reversible_rate_matrix(rng, k) if k > 1 else np.zeros((1, 1))
    pi = stationary_of(q) if k > 1 else np.ones(1)
    t = np.array([0.0, 1e-4, 0.03, 0.1, 1.0, 5.0, 40.0])
    got = thmm.expm_reversible(q, pi, t)
    want = np.stack([expm(q * x) for x in t])
    assert np.max(np.abs(got - want)) <= 1e-11
    assert np.max(np.abs(got.sum(axis=2) - 1.0)) <= 1e-12
    assert np.all(got >= 0.0)
    assert np.max(np.abs(got[0] - np.eye(k))) <= 1e-13


def test_jukes_cantor_closed_form():
    q = (np.ones((4, 4)) - 4 * np.eye(4)) / 3.0
    t = np.linspace(0.0, 3.0, 13)
    got = thmm.expm_reversible(q, np.full(4, 0.25), t)
    same = 0.25 + 0.75 * np.exp(-4.0 * t / 3.0)
    other = 0.25 - 0.25 * np.exp(-4.0 * t / 3.0)
    want = other[:, None, None] * np.ones((1, 4, 4)) + (same - other)[:, None, None] * np.eye(4)[None]
    assert np.max(np.abs(got - want)) <= 1e-14


def test_argument_checks():
    rng = np.random.default_rng(3)
    q = reversible_rate_matrix(rng, 4)
    pi = stationary_of(q)
    bad = q.copy()
    bad[0, 1] *= 1.5
    bad[0, 0] -= bad[0, 1] / 3.0
    with pytest.raises(thmm.InvalidArgument, match="not reversible"):
        thmm.expm_reversible(bad, pi, [0.1])
    with pytest.raises(thmm.InvalidArgument, match="positive"):
        thmm.expm_reversible(q, np.array([0.5, 0.5, 0.0, 0.0]), [0.1])
    with pytest.raises(thmm.InvalidArgument, match="nonnegative"):
        thmm.expm_reversible(q, pi, [0.1, -0.2])
    with pytest.raises(thmm.InvalidArgument, match="nonnegative"):
        thmm.expm_reversible(q, pi, [np.nan])
    negative = q.copy()
    negative[1, 2] = -0.1
    with pytest.raises(thmm.InvalidArgument):
        thmm.expm_reversible(negative, pi, [0.1])
