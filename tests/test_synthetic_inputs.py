"""The portable input recipe of the benchmark workloads (bench.py): SplitMix64
streams, their uniform, normal and exponential transforms, and the topology
generators, on the host and through torch (the device path runs the same
torch code on the GPU and is checked against the host at every bench run)."""

from __future__ import annotations

import math

import numpy as np

import bench


def test_splitmix64_known_answers():
    # Published first outputs of SplitMix64 for seed 1234567 (the reference
    # implementation splitmix64.c of Vigna's xoshiro page).
    expected = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                16408922859458223821]
    u = bench.uniforms_host(1234567, 0, 5)
    assert np.array_equal(u, np.array([x >> 11 for x in expected], np.float64) * 2.0**-53)


def test_torch_stream_equals_numpy_stream():
    import torch

    seed = bench.config_seed(4, bench.STREAM_EDGES)
    host = bench.uniforms_host(seed, 777, 100_000)
    dev = bench.uniforms_device(torch, seed, 777, 100_000, "cpu").numpy()
    assert np.array_equal(host, dev)
    # any range of a stream is the matching slice of the whole
    assert np.array_equal(bench.uniforms_host(seed, 0, 1000)[500:], bench.uniforms_host(seed, 500, 500))
    hn = bench.normals_host(seed, 0, 10_000)
    dn = bench.normals_device(torch, seed, 0, 10_000, "cpu").numpy()
    assert np.max(np.abs(hn - dn)) <= 1e-12


def test_distributions():
    import torch

    u = bench.uniforms_host(bench.config_seed(1), 0, 200_000)
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.002
    z = bench.normals_host(bench.config_seed(2), 0, 200_000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    rows = bench.fill_dirichlet_rows_device(torch, torch.empty((1000, 8, 8), dtype=torch.float64), 5)
    assert torch.allclose(rows.sum(2), torch.ones(1000, 8, dtype=torch.float64))
    assert float(rows.min()) > 0.0
    v = bench.variates_device(torch, 9, 100_000, device="cpu")
    assert v.dtype == torch.float32 and float(v.max()) < 1.0 and float(v.min()) >= 0.0
    assert torch.equal(v.double(), (v.double() * 2**24).floor() * 2.0**-24)


def test_forest_and_tree_generators():
    parts = bench.forest_parts(4, trees=30)
    assert len(parts) == 30
    for t, p in enumerate(parts):
        n = p.size
        assert 16 <= n <= 65536 and p[0] == -1 and np.all(p[1:] >= 0) and np.all(p[1:] < np.arange(1, n))
        if t % 3 == 2:
            assert np.all(p[1:] == 0)
    assert [p.size for p in parts] == [p.size for p in bench.forest_parts(4, trees=30)]
    sizes = np.floor(np.exp(math.log(16) + bench.uniforms_host(bench.config_seed(4), 0, 30) * math.log(65537 / 16)))
    assert np.array_equal(sizes.astype(np.int64), [p.size for p in parts])
    y = bench.yule_tree(64, bench.uniforms_host(3, 0, 63))
    assert y.size == 127 and np.all(np.bincount(y[1:], minlength=127)[np.unique(y[1:])] == 2)
    trees = bench.recursive_trees(1, 3, 50)
    assert len(trees) == 3 and all(np.all(t[1:] < np.arange(1, 50)) for t in trees)
