"""Throughput benchmark: nodes/s of log Z + adjoint (posterior marginals =
d log Z / d log potentials) on one MI355X.

Default workload (the one the headline target is quoted on): 4,096 trees
with node counts log-uniform in [16, 65,536], topologies cycling random
recursive / caterpillar / star, K = 16, FP32 form, contracted together as one
forest (batch 1, one log partition per tree). Inputs are generated and kept
resident in HBM before the timed region; one step is one full log Z +
adjoint pass over the forest. The plain FP64 form is timed on the same
inputs, and the FP64 CPU oracle is timed on a bounded sample of the trees.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 4]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "nodes/s logZ+adjoint: FP32-MFMA form vs plain FP64 at rel err <=1e-5; % roofline"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------


def recursive_tree(n, rng):
    p = np.empty(n, np.int64)
    p[0] = -1
    if n > 1:
        p[1:] = (rng.random(n - 1) * np.arange(1, n)).astype(np.int64)
    return p


def caterpillar_tree(n):
    i = np.arange(n, dtype=np.int64)
    p = np.where(i % 2 == 0, i - 2, i - 1)
    p[0] = -1
    return p


def star_tree(n):
    p = np.zeros(n, np.int64)
    p[0] = -1
    return p


def forest_workload(trees: int, low: int, high: int, seed: int, scale: float = 1.0):
    """Per-tree parent arrays of the uneven mixed-topology forest (numpy's
    generator; the benchmark itself uses forest_parts)."""
    rng = np.random.default_rng(seed)
    count = max(3, int(round(trees * scale)))
    sizes = np.floor(np.exp(rng.uniform(math.log(low), math.log(high + 1), size=count))).astype(np.int64)
    parts = []
    for t, n in enumerate(sizes):
        kind = t % 3
        parts.append(recursive_tree(n, rng) if kind == 0 else caterpillar_tree(n) if kind == 1 else star_tree(n))
    return parts


# ---------------------------------------------------------------------------
# Portable synthetic inputs: SplitMix64 (Steele, Lea and Flood, "Fast
# splittable pseudorandom number generators", OOPSLA 2014) -> uniform doubles
# (x >> 11) * 2^-53 (the conversion of tests/inference_test.cc:252-257),
# Box-Muller normals and -log(1 - u) exponentials. Output i (0-based) of a
# stream is the mix of seed + (i + 1) * gamma, so any range of a stream can
# be generated on its own, on the host (numpy uint64) or on the device
# (torch int64, whose multiplication wraps modulo 2^64 like uint64).
# The seed of stream s of config c is 0x54484D4D_0000_000c with s in bits
# 16..31 (stream 0 is the plain config seed).
# ---------------------------------------------------------------------------

_GAMMA = 0x9E3779B97F4A7C15
_MIX1 = 0xBF58476D1CE4E5B9
_MIX2 = 0x94D049BB133111EB

STREAM_TOPOLOGY, STREAM_NODES, STREAM_EDGES, STREAM_VARIATES, STREAM_ZEROS = 0, 1, 2, 3, 4
STREAM_OBSERVATIONS, STREAM_MISSING, STREAM_ROOT, STREAM_RATES, STREAM_LENGTHS = 5, 6, 7, 8, 9


def config_seed(config: int, stream: int = 0) -> int:
    return 0x54484D4D00000000 | (stream << 16) | config


def uniforms_host(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs start .. start + count - 1 of a SplitMix64 stream as doubles in [0, 1)."""
    i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = i * np.uint64(_GAMMA) + np.uint64(seed)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(_MIX1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(_MIX2)
        z ^= z >> np.uint64(31)
    return (z >> np.uint64(11)).astype(np.float64) * 2.0**-53


def _signed(x: int) -> int:
    return x - (1 << 64) if x >= 1 << 63 else x


def uniforms_device(torch, seed: int, start: int, count: int, device="cuda"):
    """uniforms_host on the device (float64 tensor), the same values."""
    z = torch.arange(start + 1, start + count + 1, dtype=torch.int64, device=device)
    z.mul_(_signed(_GAMMA)).add_(_signed(seed))

    def srl(x, s):  # logical right shift of the int64 bit pattern
        return (x >> s) & ((1 << (64 - s)) - 1)

    z ^= srl(z, 30)
    z.mul_(_signed(_MIX1))
    z ^= srl(z, 27)
    z.mul_(_signed(_MIX2))
    z ^= srl(z, 31)
    return srl(z, 11).double().mul_(2.0**-53)


def normals_host(seed: int, start: int, count: int) -> np.ndarray:
    """Box-Muller normals; normal i uses uniforms start + 2*(i//2) and +1
    (start even): r = sqrt(-2 log(1 - u0)), theta = 2 pi u1, the pair is
    (r cos theta, r sin theta)."""
    pairs = (count + 1) // 2
    u = uniforms_host(seed, start, 2 * pairs).reshape(pairs, 2)
    r = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
    t = 2.0 * math.pi * u[:, 1]
    return np.stack((r * np.cos(t), r * np.sin(t)), 1).reshape(-1)[:count]


def normals_device(torch, seed: int, start: int, count: int, device="cuda"):
    pairs = (count + 1) // 2
    u = uniforms_device(torch, seed, start, 2 * pairs, device).view(pairs, 2)
    r = torch.sqrt(-2.0 * torch.log1p(-u[:, 0]))
    t = (2.0 * math.pi) * u[:, 1]
    return torch.stack((r * torch.cos(t), r * torch.sin(t)), 1).view(-1)[:count]


def check_device_stream(torch, device="cuda") -> None:
    """The device generator reproduces the host one (bit patterns of the
    uniforms and normals within rounding of the device's libm)."""
    seed = config_seed(4, STREAM_EDGES)
    host = uniforms_host(seed, 12345, 1 << 16)
    dev = uniforms_device(torch, seed, 12345, 1 << 16, device).cpu().numpy()
    if not np.array_equal(host, dev):
        raise RuntimeError("device SplitMix64 stream differs from the host stream")
    hn = normals_host(seed, 0, 1 << 12)
    dn = normals_device(torch, seed, 0, 1 << 12, device).cpu().numpy()
    if np.max(np.abs(hn - dn)) > 1e-12:
        raise RuntimeError("device Box-Muller normals differ from the host ones")


def variates_device(torch, seed: int, count: int, dtype=None, device="cuda"):
    """Sampling variates: uniforms truncated to 24 bits, floor(u 2^24) 2^-24,
    which FP32 holds exactly, so every variate stays below 1 in FP32 and in
    FP64 alike."""
    u = uniforms_device(torch, seed, 0, count, device).mul_(2.0**24).floor_().mul_(2.0**-24)
    return u.to(dtype or torch.float32)


_CHUNK = 1 << 26  # values per device generation pass (even)


def fill_lognormal_device(torch, out, seed: int, sigma: float = 1.0, log: bool = False):
    """out.view(-1)[i] = exp(sigma * normal_i) (or sigma * normal_i with
    log=True), rounded once from FP64 to out's type."""
    flat = out.view(-1)
    for i in range(0, flat.numel(), _CHUNK):
        n = min(_CHUNK, flat.numel() - i)
        v = normals_device(torch, seed, i, n, flat.device).mul_(sigma)
        flat[i:i + n] = v if log else v.exp_()
        del v
    return out


def fill_dirichlet_rows_device(torch, out, seed: int):
    """Rows of the last dimension ~ Dirichlet(1): Exp(1) = -log(1 - u) draws
    normalised by their row sum in FP64, rounded once to out's type."""
    k = out.shape[-1]
    rows = out.view(-1, k)
    per = max(1, _CHUNK // k)
    for r in range(0, rows.shape[0], per):
        m = min(per, rows.shape[0] - r)
        v = uniforms_device(torch, seed, r * k, m * k, rows.device).view(m, k)
        v = torch.log1p(v.neg_()).neg_()
        rows[r:r + m] = v.div_(v.sum(1, keepdim=True))
        del v
    return out


def forest_parts(config: int = 4, trees: int = 4096, low: int = 16, high: int = 65536, scale: float = 1.0):
    """The uneven forest from SplitMix64 stream 0 of the config: sizes
    floor(exp(ln low + u (ln(high + 1) - ln low))) from uniforms 0 .. T-1,
    topology t % 3 (random recursive, caterpillar, star), and the random
    recursive trees' parents parent(i) = floor(u i) from the uniforms that
    follow, tree after tree."""
    seed = config_seed(config, STREAM_TOPOLOGY)
    count = max(3, int(round(trees * scale)))
    u = uniforms_host(seed, 0, count)
    sizes = np.floor(np.exp(math.log(low) + u * (math.log(high + 1) - math.log(low)))).astype(np.int64)
    parts, cursor = [], count
    for t, n in enumerate(sizes):
        kind = t % 3
        if kind == 0:
            p = np.empty(n, np.int64)
            p[0] = -1
            if n > 1:
                p[1:] = np.floor(uniforms_host(seed, cursor, n - 1) * np.arange(1, n)).astype(np.int64)
            cursor += n - 1
            parts.append(p)
        else:
            parts.append(caterpillar_tree(n) if kind == 1 else star_tree(n))
    return parts


def recursive_trees(config: int, trees: int, n: int):
    """`trees` random recursive trees of n nodes from stream 0 of the config."""
    seed = config_seed(config, STREAM_TOPOLOGY)
    u = uniforms_host(seed, 0, trees * (n - 1)).reshape(trees, n - 1)
    out = []
    for t in range(trees):
        p = np.empty(n, np.int64)
        p[0] = -1
        p[1:] = np.floor(u[t] * np.arange(1, n)).astype(np.int64)
        out.append(p)
    return out


def yule_tree(leaves: int, uniforms):
    """Random full binary tree: split a uniform random leaf until `leaves`;
    split j takes leaf floor(u_j * current leaves) of the list of leaves."""
    parents = np.empty(2 * leaves - 1, np.int64)
    parents[0] = -1
    current = [0]
    size = 1
    for u in uniforms[: leaves - 1].tolist():
        i = int(u * len(current))
        leaf = current[i]
        parents[size] = leaf
        parents[size + 1] = leaf
        current[i] = size
        current.append(size + 1)
        size += 2
    return parents


def concatenate(parts):
    offsets = np.zeros(len(parts) + 1, np.int64)
    offsets[1:] = np.cumsum([p.size for p in parts])
    out = np.empty(offsets[-1], np.int64)
    for t, p in enumerate(parts):
        out[offsets[t] : offsets[t + 1]] = np.where(p < 0, -1, p + offsets[t])
    return out, offsets


# ---------------------------------------------------------------------------
# Algorithmic bytes per op (FP32 s = 4; per-item edge slots at batch 1).
# Values only; the FP64 scale words (one per vector, K per matrix) are left
# out. Forward: rake reads the leaf vector and the K x K slot and writes the
# branch; combination/absorption read two vectors and write one; a fused
# rake level (the rakes, combinations and absorptions of a round in one
# launch) counts as its rakes, its partial-sum levels as the FP64 partials
# they read; compression reads L, R and the middle vector and writes the
# composed path. Reverse:
# rake reads the slot, the leaf vector and the branch adjoint and writes the
# slot adjoint (an edge marginal) and the leaf adjoint; combination and
# absorption copy one adjoint vector; compression reads L, R, the middle
# vector and the path adjoint and writes the adjoints of L, R and the middle.
# A grouped compression of n slots (primitive 6) reads the n slots and the
# n - 1 middle vectors and writes one path; its reverse reads the slots, the
# middles and the path adjoint and writes n slot adjoints and n - 1 middle
# adjoints. Its levels report these as K-vectors moved per item (reverse
# levels split into groups of 2 slots, primitive 8, of 3 or 4, primitive 7,
# and the rest).
# ---------------------------------------------------------------------------


def op_bytes(phase: int, primitive: int, k: int, s: int) -> int:
    kk = k * k
    if phase == 1:  # forward (4: fused rake level, per rake; 5: partial sums, per FP64 partial read)
        return {0: (2 * k + kk) * s, 1: 3 * k * s, 2: 3 * k * s, 3: (3 * kk + k) * s, 4: (2 * k + kk) * s,
                5: 8 * k, 6: k * s, 9: 0}[primitive]
    if phase == 3:  # reverse
        return {0: (2 * kk + 3 * k) * s, 1: 2 * k * s, 2: 2 * k * s, 3: (5 * kk + 2 * k) * s, 6: k * s,
                7: k * s, 8: k * s, 9: 0}[primitive]
    return 0


KERNEL_NAMES = {
    (1, 0): "RakeForward",
    (1, 1): "CombineForward",
    (1, 2): "AbsorbForward",
    (1, 3): "CompressForward",
    (1, 4): "RakeGather",
    (1, 5): "GatherSegments",
    (1, 6): "ChainForward",
    (3, 0): "RakeReverse",
    (3, 1): "CombineReverse",
    (3, 2): "AbsorbReverse",
    (3, 3): "CompressReverse",
    (3, 6): "ChainReverse",
    (3, 7): "ChainReverseSmall",
    (3, 8): "ChainReversePair",
    (1, 9): "PersistentForward",
    (3, 9): "PersistentReverse",
}


# ---------------------------------------------------------------------------
# GPU leg
# ---------------------------------------------------------------------------


def run_gpu(thmm, torch, plan, k, precision, nodes, edges, steps, warmup, tune=True, half_edges=False, use=None):
    """One device-resident marginal pass per step; `use` installs given
    launch variants instead of autotuning."""
    dtype = torch.float32 if precision == "f32" else torch.float64
    nodes = nodes.to(dtype)
    edges = edges.to(dtype)
    n, e = plan.num_nodes, plan.num_edges
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, 1, precision=precision, device_resident=True, half_edge_marginals=half_edges)
    log_z = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    node_out = torch.empty(n * k, dtype=dtype, device="cuda")
    edge_out = torch.empty(e * k * k, dtype=torch.float16 if half_edges else dtype, device="cuda")
    stream = torch.cuda.current_stream()
    args = (nodes.data_ptr(), edges.data_ptr(), 1, log_z.data_ptr(), node_out.data_ptr(), edge_out.data_ptr(), stream.cuda_stream)
    for _ in range(warmup):
        ws.marginals_device(*args)
    torch.cuda.synchronize()
    # Launch variants picked by the library's autotuner on these inputs
    # (outside the timed region).
    if use is not None:
        ws.set_tuning(**{f: v for f, v in use.items() if f != "ms"})
        ws.marginals_device(*args)
        tuning = dict(ws.tuning(), ms=None)
    else:
        tuning = ws.autotune_marginals_device(*args) if tune else dict(ws.tuning(), ms=None)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        ws.marginals_device(*args)
    stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    # Per-kernel figures come from the level launches (a persistent pass is
    # one launch per direction).
    chosen = ws.tuning()
    if chosen.get("persistent", 0):
        ws.set_tuning(persistent=0)
    timings = ws.profile_marginals_device(*args)
    if chosen.get("persistent", 0):
        ws.set_tuning(persistent=1)
    torch.cuda.synchronize()
    return ws, ms, timings, log_z, node_out, edge_out, tuning


def run_map_and_sampling(thmm, torch, plan, k, nodes, edges, args, variate_seed):
    """MAP (thmm_maximum_device) and posterior sampling (thmm_sample_device,
    one U[0,1) variate per node) on the same resident FP32 inputs, each in
    its own device-resident workspace; both run in FP64 internally. Times
    `steps` calls after `warmup`, and checks that repeated calls give the
    same states and that sampled states lie in range."""
    n = plan.num_nodes
    roots = plan.roots.size
    out = {}
    stream = torch.cuda.current_stream()
    for name, flag in (("map", thmm._abi.THMM_RESERVE_MAXIMUM), ("sample", thmm._abi.THMM_RESERVE_SAMPLING)):
        ws = thmm.Workspace()
        ws.reserve(plan, k, 1, precision="f32", flags=flag | thmm._abi.THMM_RESERVE_DEVICE_RESIDENT)
        log_z = torch.empty(roots, dtype=torch.float64, device="cuda")
        states = torch.empty(n, dtype=torch.int32, device="cuda")
        if name == "map":
            call = lambda: ws.maximum_device(nodes.data_ptr(), edges.data_ptr(), 1, log_z.data_ptr(),  # noqa: E731
                                             states.data_ptr(), stream.cuda_stream)
        else:
            uniforms = variates_device(torch, variate_seed, n)
            failure = torch.zeros(1, dtype=torch.int32, device="cuda")
            call = lambda: ws.sample_device(nodes.data_ptr(), edges.data_ptr(), uniforms.data_ptr(), 1,  # noqa: E731
                                            log_z.data_ptr(), states.data_ptr(), failure.data_ptr(),
                                            stream.cuda_stream)
        for _ in range(max(1, args.warmup)):
            call()
        torch.cuda.synchronize()
        first = states.clone()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(args.steps):
            call()
        stop.record(stream)
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / args.steps
        entry = {"ms_per_step": round(ms, 4), "value": n / (ms * 1e-3), "unit": "nodes/s",
                 "repeatable": bool(torch.equal(first, states)),
                 "states_in_range": bool(((states >= 0) & (states < k)).all()),
                 "log_z_finite": bool(torch.isfinite(log_z).all())}
        if name == "sample":
            entry["failure_flag"] = int(failure.item())
        out[name] = entry
        ws.close()
        del ws, log_z, states
        torch.cuda.empty_cache()
    return out


def measured_traffic(workload, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summaries (FETCH_SIZE + WRITE_SIZE, scaled by a copy of
    known size), or None when no summary covers it."""
    for path in sorted((ROOT / "profiles").glob("traffic_*.md")):
        for line in path.read_text().splitlines():
            parts = line.split()
            if len(parts) == 4 and parts[0] == "traffic_bytes_per_launch" and parts[1:3] == [workload, kernel]:
                return int(parts[3])
    return None


class SharedEdgeReads:
    """For a batch of items that share one set of edge potentials: per plan
    level, how many of its edge-slot reads name a caller edge (read once per
    op for the whole batch) rather than a composed path (one per item). A
    compression's path replaces its left edge's slot (src/inference.cc:
    191-214), which is replayed here."""

    def __init__(self, plan):
        e = plan.num_edges
        current = np.arange(e, dtype=np.int64)  # < e: the caller's edge
        rounds = plan.round_batch_offsets
        self.rake_rounds, self.rake_levels, self.compress_levels = [], [], []
        for r in range(plan.num_rounds):
            rakes = shared = 0
            for i in range(int(rounds[r]), int(rounds[r + 1])):
                b = plan.primitive_batches[i]
                lo, hi = int(b["offset"]), int(b["offset"]) + int(b["count"])
                if b["primitive"] == 0:
                    sh = int(np.count_nonzero(current[plan.rakes["edge"][lo:hi].astype(np.int64)] < e))
                    self.rake_levels.append((hi - lo, sh))
                    rakes, shared = rakes + hi - lo, shared + sh
                elif b["primitive"] == 3:
                    ops = plan.compressions[lo:hi]
                    sh = int(np.count_nonzero(current[ops["left_edge"].astype(np.int64)] < e)) + int(
                        np.count_nonzero(current[ops["right_edge"].astype(np.int64)] < e))
                    self.compress_levels.append((hi - lo, sh))
                    current[ops["left_edge"].astype(np.int64)] = e + np.arange(lo, hi)
            if rakes:
                self.rake_rounds.append((rakes, shared))

    def bytes_of(self, timings, k, s):
        """Algorithmic bytes of each profiled launch, in order."""
        kk = k * k
        rounds, rakes, compresses = iter(self.rake_rounds), iter(self.rake_levels), iter(self.compress_levels)
        back_rakes, back_compresses = iter(self.rake_levels[::-1]), iter(self.compress_levels[::-1])
        out = []
        for phase, primitive, count, batch, _ in timings:
            if phase == 1 and primitive in (0, 4):
                c, sh = next(rounds if primitive == 4 else rakes)
                out.append(batch * c * 2 * k * s + (sh + (c - sh) * batch) * kk * s)
            elif phase == 1 and primitive == 3:
                c, sh = next(compresses)
                out.append(batch * c * (kk + k) * s + (sh + (2 * c - sh) * batch) * kk * s)
            elif phase == 3 and primitive == 0:
                c, sh = next(back_rakes)
                out.append(batch * c * (kk + 3 * k) * s + (sh + (c - sh) * batch) * kk * s)
            elif phase == 3 and primitive == 3:
                c, sh = next(back_compresses)
                out.append(batch * c * (3 * kk + 2 * k) * s + (sh + (2 * c - sh) * batch) * kk * s)
            elif phase in (1, 3):
                out.append(op_bytes(phase, primitive, k, s) * count * batch)
            else:
                out.append(0)
        return out


def roofline_from(timings, k, s, reads=None, workload=None):
    """The profiled kernel family with the most time, its algorithmic GB/s
    and the breakdown; `reads` (SharedEdgeReads) counts edges shared by a
    batch once per op."""
    totals = {}
    nbytes_of = reads.bytes_of(timings, k, s) if reads is not None else [
        op_bytes(ph, pr, k, s) * c * b if ph in (1, 3) else 0 for ph, pr, c, b, _ in timings]
    for (phase, primitive, count, batch, ms), nb in zip(timings, nbytes_of):
        if phase not in (1, 3):
            continue
        key = (phase, primitive)
        t = totals.setdefault(key, [0.0, 0, 0])
        t[0] += ms
        t[1] += nb
        t[2] += 1
    key = max(totals, key=lambda x: totals[x][0])
    ms, nbytes, launches = totals[key]
    achieved = nbytes / (ms * 1e-3) / 1e9
    breakdown = {KERNEL_NAMES[kk]: round(v[0], 4) for kk, v in sorted(totals.items(), key=lambda x: -x[1][0])}
    gbs = {KERNEL_NAMES[kk]: round(v[1] / (v[0] * 1e-3) / 1e9, 1) for kk, v in sorted(totals.items(), key=lambda x: -x[1][0])
           if v[0] > 0 and v[1] > 0}
    per_launch = {KERNEL_NAMES[kk]: int(v[1] / v[2]) for kk, v in sorted(totals.items(), key=lambda x: -x[1][0]) if v[1] > 0}
    return {
        "kernel": KERNEL_NAMES[key],
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(workload, KERNEL_NAMES[key]),
        "launches": launches,
        "avg_launch_us": round(ms / launches * 1e3, 3),
        "algorithmic_bytes_per_launch": int(nbytes / launches),
        "kernel_ms_breakdown": breakdown,
        "kernel_algorithmic_gbs": gbs,
        "kernel_algorithmic_bytes_per_launch": per_launch,
    }


# Work items a block of each profiled launch kind covers at batch 1: the
# team rake kernels hold 16 teams per 256-lane block (K <= 16), the grouped
# compression kernels 4 groups (one wavefront each).
ITEMS_PER_BLOCK = {(1, 4): 16, (1, 5): 16, (1, 6): 4, (3, 0): 16, (3, 6): 4, (3, 7): 4, (3, 8): 4, (1, 3): 8, (3, 3): 8}


def level_table(timings, k, s, path, title, ops_of=None):
    """Markdown table of every profiled launch of one marginal pass: kind,
    op count, microseconds, algorithmic GB/s and blocks; totals the time of
    launches with fewer than 256 blocks (too few to fill 256 CUs)."""
    rows, small_ms, total_ms = [], 0.0, 0.0
    for i, (phase, primitive, count, batch, ms) in enumerate(timings):
        total_ms += ms
        if phase not in (1, 3):
            rows.append(f"| {i} | {'staging' if phase == 0 else 'finish + seed' if phase == 2 else 'output'} | — | — | "
                        f"{ms * 1e3:.1f} | — | — |")
            continue
        nbytes = op_bytes(phase, primitive, k, s) * count * batch
        ops = ops_of[i] if ops_of else count
        per = ITEMS_PER_BLOCK.get((phase, primitive))
        blocks = -(-ops * batch // per) if per else None
        if blocks is not None and blocks < 256:
            small_ms += ms
        rows.append(f"| {i} | {KERNEL_NAMES[(phase, primitive)]} | {ops} | {count} | {ms * 1e3:.1f} | "
                    f"{nbytes / (ms * 1e-3) / 1e9:.0f} | {blocks if blocks is not None else '—'} |")
    text = [f"# {title}", "",
            f"One marginal pass, {len(timings)} launches, {total_ms:.3f} ms in total; launches with fewer than 256 "
            f"blocks (too few to fill 256 CUs): {small_ms:.3f} ms ({small_ms / total_ms:.1%}).", "",
            "`ops` counts the level's work units (rakes, partial sums or compression groups), `count` the units the "
            "algorithmic bytes are counted in (K-vectors for grouped compressions).", "",
            "| # | kernel | ops | count | us | algorithmic GB/s | blocks |", "|---:|---|---:|---:|---:|---:|---:|"] + rows
    Path(path).write_text("\n".join(text) + "\n")
    return small_ms, total_ms


# ---------------------------------------------------------------------------
# CPU baseline + parity sample (oracle)
# ---------------------------------------------------------------------------


def tree_slices(offsets, t):
    lo, hi = int(offsets[t]), int(offsets[t + 1])
    return lo, hi, lo - t, hi - t - 1  # node range, edge range


def tree_inputs(nodes, edges, offsets, t):
    """FP64 host copies of one tree's inputs (sliced on the device)."""
    lo, hi, elo, ehi = tree_slices(offsets, t)
    return nodes[lo:hi].double().cpu().numpy(), edges[elo:ehi].double().cpu().numpy()


def host_cpus():
    """The host's CPUs: what this process may use (its affinity mask), what
    the machine reports, and the model name."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        import subprocess

        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except Exception:
        nproc = usable
    return {"nproc": nproc, "os_cpu_count": os.cpu_count(), "affinity": usable, "model": model}


# One GPU's share of a GPU box's host is 16 CPUs, whatever the machine
# reports; the CPU leg uses that many threads (fewer if the process may use
# fewer).
CPU_SHARE = 16


def timed_median(fn, repeats: int = 5):
    """One warm-up call, then the median wall time of `repeats` calls."""
    fn()
    times = []
    for _ in range(repeats):
        start = time.perf_counter()
        fn()
        times.append(time.perf_counter() - start)
    return float(np.median(times)), times


def cpu_baseline(thmm, oracle, parts, offsets, nodes, edges, k, budget_nodes, repeats=5):
    """Times the FP64 oracle (log Z + node and edge marginals) on a bounded
    sample of whole trees, one tree per task over CPU_SHARE host threads:
    one warm-up pass, then the median of `repeats` passes (the method of
    benchmarks/inference_benchmark.cc:19-36,86-121)."""
    host = host_cpus()
    threads = max(1, min(CPU_SHARE, host["affinity"]))
    order = np.random.default_rng(7).permutation(len(parts))
    chosen, total = [], 0
    for t in order:
        if total >= budget_nodes:
            break
        if parts[t].size > budget_nodes // 2:
            continue
        chosen.append(int(t))
        total += parts[t].size
    plans = {t: thmm.Plan(parts[t]) for t in chosen}
    jobs = [(t, *tree_inputs(nodes, edges, offsets, t)) for t in chosen]

    def work(job):
        t, nn, ee = job
        return oracle.marginals(plans[t], k, nn, ee)

    with ThreadPoolExecutor(max_workers=threads) as pool:
        seconds, times = timed_median(lambda: list(pool.map(work, jobs)), repeats)
    return {"value": total / seconds, "unit": "nodes/s", "cores": threads, "kind": "port",
            "sample": f"{len(chosen)} whole trees of the forest, {total} nodes, FP64 oracle log Z + node and edge "
                      f"marginals; median of {repeats} passes after one warm-up ({seconds:.2f}s per pass)",
            "host": host, "pass_seconds": [round(x, 3) for x in times]}


def parity_sample(thmm, oracle, parts, offsets, nodes, edges, k, log_z, node_out, edge_out, picks):
    """Worst errors of the GPU outputs against the oracle on a few trees."""
    worst = {"log_z_rel": 0.0, "node_abs": 0.0, "edge_abs": 0.0}
    for t in picks:
        plan = thmm.Plan(parts[t])
        lo, hi, elo, ehi = tree_slices(offsets, t)
        nn, ee = tree_inputs(nodes, edges, offsets, t)
        lz, nm, em = oracle.marginals(plan, k, nn, ee)
        worst["log_z_rel"] = max(worst["log_z_rel"], abs(log_z[t] - lz[0]) / max(1.0, abs(lz[0])))
        gpu_nodes = node_out[lo:hi].double().cpu().numpy()
        worst["node_abs"] = max(worst["node_abs"], float(np.max(np.abs(gpu_nodes - nm))))
        if ehi > elo:
            gpu_edges = edge_out[elo:ehi].double().cpu().numpy()
            worst["edge_abs"] = max(worst["edge_abs"], float(np.max(np.abs(gpu_edges - em))))
    return worst


# ---------------------------------------------------------------------------


def time_calls(torch, call, steps, warmup):
    """Event-timed milliseconds per call of `call` on torch's current stream
    (warm-up calls first, untimed)."""
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        call()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def marginal_pass(thmm, torch, plan, k, batch, precision, nodes, edges, graphs=False):
    """A device-resident marginal workspace, its outputs and a call."""
    dtype = torch.float32 if precision == "f32" else torch.float64
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, batch, precision=precision, device_resident=True, graphs=graphs)
    out = {"log_z": torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda"),
           "nodes": torch.empty((batch, plan.num_nodes, k), dtype=dtype, device="cuda"),
           "edges": torch.empty((batch, plan.num_edges, k, k), dtype=dtype, device="cuda")}
    stream = torch.cuda.current_stream().cuda_stream
    args = (nodes.data_ptr(), edges.data_ptr(), batch, out["log_z"].data_ptr(), out["nodes"].data_ptr(),
            out["edges"].data_ptr(), stream)
    return ws, out, (lambda: ws.marginals_device(*args)), args


def result_line(metric, value, args, ms, dtype, data, config, **extra):
    line = {"metric": metric, "value": value, "unit": "nodes/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": data, "config": config}
    line.update(extra)
    print(json.dumps(line), flush=True)
    return 0


def cpu_median(label, total_nodes, fn, repeats=5):
    host = host_cpus()
    seconds, times = timed_median(fn, repeats)
    return {"value": total_nodes / seconds, "unit": "nodes/s", "cores": max(1, min(CPU_SHARE, host["affinity"])),
            "kind": "port", "sample": f"{label}; median of {repeats} passes after one warm-up ({seconds:.3f}s per pass)",
            "host": host, "pass_seconds": [round(x, 4) for x in times]}


def run_recursive_fp64(args, thmm, oracle, torch) -> int:
    """64 random recursive trees of 1,023 nodes (parent(i) ~ U{0..i-1}), K = 4,
    plain FP64, as one forest (per-tree edges): log Z + node and edge
    marginals per step, replayed as a captured hipGraph."""
    k, trees, n_tree = 4, 64, 1023
    parts = recursive_trees(1, trees, n_tree)
    parents, offsets = concatenate(parts)
    plan = thmm.Plan(parents)
    n, e = plan.num_nodes, plan.num_edges
    check_device_stream(torch)
    nodes = fill_lognormal_device(torch, torch.empty((1, n, k), dtype=torch.float64, device="cuda"),
                                  config_seed(1, STREAM_NODES))
    edges = fill_dirichlet_rows_device(torch, torch.empty((e, k, k), dtype=torch.float64, device="cuda"),
                                       config_seed(1, STREAM_EDGES))
    ws, out, call, margs = marginal_pass(thmm, torch, plan, k, 1, "f64", nodes, edges, graphs=True)
    ms = time_calls(torch, call, args.steps, args.warmup)
    wse, oute, calle, eargs = marginal_pass(thmm, torch, plan, k, 1, "f64", nodes, edges)
    ms_eager = time_calls(torch, calle, args.steps, args.warmup)
    timings = wse.profile_marginals_device(*eargs)
    roofline = roofline_from(timings, k, 8, workload="recursive64_k4")
    # Parity: all 64 trees against the FP64 oracle.
    lz = out["log_z"].cpu().numpy()
    nn, ee = nodes[0].cpu().numpy(), edges.cpu().numpy()
    gn, ge = out["nodes"][0].cpu().numpy(), out["edges"][0].cpu().numpy()
    tree_plans = [thmm.Plan(p) for p in parts]
    worst = {"log_z_rel": 0.0, "node_abs": 0.0, "edge_abs": 0.0}

    def tree_job(t):
        lo, hi, elo, ehi = tree_slices(offsets, t)
        return t, oracle.marginals(tree_plans[t], k, nn[lo:hi], ee[elo:ehi])

    with ThreadPoolExecutor(max_workers=max(1, min(CPU_SHARE, host_cpus()["affinity"]))) as pool:
        for t, (olz, onm, oem) in pool.map(tree_job, range(trees)):
            lo, hi, elo, ehi = tree_slices(offsets, t)
            worst["log_z_rel"] = max(worst["log_z_rel"], abs(lz[t] - olz[0]) / max(1.0, abs(olz[0])))
            worst["node_abs"] = max(worst["node_abs"], float(np.max(np.abs(gn[lo:hi] - onm))))
            worst["edge_abs"] = max(worst["edge_abs"], float(np.max(np.abs(ge[elo:ehi] - oem))))
        cpu = None
        if not args.no_cpu:
            cpu = cpu_median(f"all {trees} trees ({n} nodes), FP64 oracle log Z + node and edge marginals, one tree "
                             "per task", n, lambda: list(pool.map(tree_job, range(trees))))
    worst["trees_checked"] = trees
    log(f"K=4 FP64 forest: graph replay {ms:.4f} ms, eager {ms_eager:.4f} ms; parity {worst}")
    ws.close()
    wse.close()
    return result_line(
        "nodes/s log Z + adjoint, plain FP64 (64 random recursive trees)", n / (ms * 1e-3), args, ms, "f64",
        "synthetic (SplitMix64 / Box-Muller, seed 0x54484D4D00000001: parent(i) = floor(u i), node potentials "
        "exp(N(0,1)), edge rows Dirichlet(1))",
        {"workload": "64 random recursive trees of 1023 nodes, K=4, plain FP64, one forest (per-tree edges), "
         "log Z + node and edge marginals, hipGraph replay", "nodes": int(n), "levels": int(plan.primitive_batches.size)},
        roofline=roofline, cpu_baseline=cpu, eager_ms_per_step=round(ms_eager, 4), parity=worst)


def run_long_chains(args, thmm, oracle, torch) -> int:
    """16 path graphs of 100,000 nodes, K = 8, FP32, one topology with 16
    batch items sharing the edges: node potentials exp(N(0, 2)) with each
    entry set to 0 with probability 0.1 (a node whose entries all drew 0
    keeps its first state at 1), edges exp(N(0, 1)). A step is log Z + node
    and edge marginals; MAP and posterior sampling (one variate per node)
    are timed on the same inputs."""
    k, batch, n_path = 8, 16, 100_000
    plan = thmm.Plan(np.arange(-1, n_path - 1, dtype=np.int64))
    n, e = plan.num_nodes, plan.num_edges
    check_device_stream(torch)
    nodes = fill_lognormal_device(torch, torch.empty((batch, n, k), device="cuda"), config_seed(2, STREAM_NODES),
                                  sigma=2.0)
    zero = uniforms_device(torch, config_seed(2, STREAM_ZEROS), 0, batch * n * k).view(batch, n, k) < 0.1
    nodes[zero] = 0.0
    dead = (nodes == 0).all(dim=2)
    nodes[..., 0][dead] = 1.0
    del zero
    edges = fill_lognormal_device(torch, torch.empty((e, k, k), device="cuda"), config_seed(2, STREAM_EDGES))
    ws, out, call, margs = marginal_pass(thmm, torch, plan, k, batch, "f32", nodes, edges)
    ms = time_calls(torch, call, args.steps, args.warmup)
    timings = ws.profile_marginals_device(*margs)
    roofline = roofline_from(timings, k, 4, SharedEdgeReads(plan), workload="chains16_k8")
    lz = out["log_z"].cpu().numpy()
    # Parity: two whole items against the FP64 oracle; every item's
    # normalisation.
    worst = {"log_z_rel": 0.0, "node_abs": 0.0, "edge_abs": 0.0}
    ee = edges.double().cpu().numpy()
    for b in (0, batch - 1):
        olz, onm, oem = oracle.marginals(plan, k, nodes[b].double().cpu().numpy(), ee)
        worst["log_z_rel"] = max(worst["log_z_rel"], abs(lz[b] - olz[0]) / max(1.0, abs(olz[0])))
        worst["node_abs"] = max(worst["node_abs"], float(np.max(np.abs(out["nodes"][b].double().cpu().numpy() - onm))))
        worst["edge_abs"] = max(worst["edge_abs"], float(np.max(np.abs(out["edges"][b].double().cpu().numpy() - oem))))
    worst["node_normalisation_abs"] = float((out["nodes"].double().sum(2) - 1.0).abs().max())
    worst["items_checked"] = 2
    ws.close()
    del ws, out
    torch.cuda.empty_cache()
    # MAP and sampling on the same inputs (device entry points, FP64 inside).
    decode = {}
    stream = torch.cuda.current_stream().cuda_stream
    uniforms = variates_device(torch, config_seed(2, STREAM_VARIATES), batch * n).view(batch, n)
    for name, flag in (("map", thmm._abi.THMM_RESERVE_MAXIMUM), ("sample", thmm._abi.THMM_RESERVE_SAMPLING)):
        wd = thmm.Workspace()
        wd.reserve(plan, k, batch, precision="f32", flags=flag | thmm._abi.THMM_RESERVE_DEVICE_RESIDENT)
        lzd = torch.empty(batch, dtype=torch.float64, device="cuda")
        states = torch.empty((batch, n), dtype=torch.int32, device="cuda")
        failure = torch.zeros(1, dtype=torch.int32, device="cuda")
        if name == "map":
            dcall = lambda: wd.maximum_device(nodes.data_ptr(), edges.data_ptr(), batch, lzd.data_ptr(),  # noqa: E731
                                              states.data_ptr(), stream)
        else:
            dcall = lambda: wd.sample_device(nodes.data_ptr(), edges.data_ptr(), uniforms.data_ptr(), batch,  # noqa: E731
                                             lzd.data_ptr(), states.data_ptr(), failure.data_ptr(), stream)
        dms = time_calls(torch, dcall, args.steps, args.warmup)
        got = states[0].cpu().numpy().astype(np.uint32)
        nn0 = nodes[0].double().cpu().numpy()
        if name == "map":
            _, expected = oracle.maximum(plan, k, nn0, ee)
            exact = bool(np.array_equal(got, expected))
        else:
            expected, dist = oracle.sample(plan, k, nn0, ee, uniforms[0].double().cpu().numpy())
            exact = bool(np.array_equal(got, expected)) if np.min(dist) >= 1e-12 else "near a CDF boundary"
        decode[name] = {"ms_per_step": round(dms, 4), "value": batch * n / (dms * 1e-3), "unit": "nodes/s",
                        "item0_bit_exact_vs_oracle": exact, "failure_flag": int(failure.item())}
        wd.close()
        del wd
        torch.cuda.empty_cache()
    log(f"K=8 chains: marginals {ms:.3f} ms; {decode}; parity {worst}")
    cpu = None
    if not args.no_cpu:
        host_nodes = nodes.double().cpu().numpy()
        threads = max(1, min(CPU_SHARE, host_cpus()["affinity"]))
        cpu = cpu_median(f"all {batch} items ({batch * n} nodes), FP64 oracle log Z + node and edge marginals, "
                         f"one item per thread", batch * n,
                         lambda: oracle.marginals_batched(plan, k, host_nodes, ee, threads=threads))
    return result_line(
        "nodes/s log Z + adjoint, FP32 (16 long chains)", batch * n / (ms * 1e-3), args, ms, "f32",
        "synthetic (SplitMix64 / Box-Muller, seed 0x54484D4D00000002: node potentials exp(N(0,2)) with 10% of "
        "entries 0, edges exp(N(0,1)), one variate per node)",
        {"workload": "16 path graphs of 100000 nodes (one topology, batch 16, shared edges), K=8, FP32 log Z + node "
         "and edge marginals", "nodes": int(batch * n), "levels": int(plan.primitive_batches.size)},
        roofline=roofline, cpu_baseline=cpu, map_and_sampling=decode, parity=worst)


def phylo_inputs(sites: int, depth: int = 12):
    """The categorical model of config 3 from its SplitMix64 streams:
    balanced binary tree of 2^(depth+1) - 1 nodes, leaves observed;
    observations floor(4 u), coded 4 (missing, an all-ones emission row)
    where a second uniform is below 0.05; root ~ Dirichlet(1); emission
    [I_4; 1^T]; Q = S Pi with S_ij ~ Exp(1) symmetric, pi ~ Dirichlet(1),
    rows summing to 0, unit mean rate; t_e ~ Exp(mean 0.1)."""
    k = 4
    n = 2 ** (depth + 1) - 1
    parents = np.arange(n, dtype=np.int64) - 1
    parents //= 2
    parents[0] = -1
    leaves = np.arange(n // 2, n, dtype=np.uint32)
    u = uniforms_host(config_seed(3, STREAM_OBSERVATIONS), 0, sites * leaves.size).reshape(sites, leaves.size)
    obs = np.floor(u * k).astype(np.uint8)
    obs[uniforms_host(config_seed(3, STREAM_MISSING), 0, obs.size).reshape(obs.shape) < 0.05] = k
    r = -np.log1p(-uniforms_host(config_seed(3, STREAM_ROOT), 0, k))
    root = r / r.sum()
    emission = np.vstack([np.eye(k), np.ones((1, k))])
    draws = -np.log1p(-uniforms_host(config_seed(3, STREAM_RATES), 0, k * k + k))
    sym = draws[: k * k].reshape(k, k)
    sym = np.triu(sym, 1)
    sym = sym + sym.T
    pi = draws[k * k:] / draws[k * k:].sum()
    q = sym * pi[None, :]
    np.fill_diagonal(q, -q.sum(axis=1))
    q /= -np.sum(pi * np.diag(q))
    t = -0.1 * np.log1p(-uniforms_host(config_seed(3, STREAM_LENGTHS), 0, n - 1))
    return parents, leaves, obs, root, emission, q, pi, t


def run_phylogenetic(args, thmm, oracle, torch) -> int:
    """1,024 sites on one balanced binary tree of 4,096 leaves (8,191 nodes),
    K = 4, FP32 categorical model. A step sets the branch lengths (edge
    factors expm(Q t_e) formed on the device), runs the forward half with the
    observations and factors resident (kNone) and the backward half giving
    d (sum of log Z) / d t_e. The kFactors form (host expm, edges restaged)
    is timed too."""
    sites, k, categories = 1024, 4, 5
    parents, leaves, obs, root, emission, q, pi, t = phylo_inputs(sites)
    plan = thmm.Plan(parents)
    n, e = plan.num_nodes, plan.num_edges
    factors = thmm.expm_reversible(q, pi, t)
    model = thmm.BatchedCategoricalModelView(plan, k, sites, categories, leaves, obs, root.astype(np.float32),
                                             emission.astype(np.float32), factors.astype(np.float32))
    ws = thmm.Workspace()
    ws.reserve_categorical_marginals(plan, k, sites, categories, leaves, precision="f32")
    ws.categorical_logz_forward(model)  # stages everything (kAll)
    ws.branch_length_backward(q, None)
    lengths = [t * (1.0 + 0.01 * i) for i in range(max(args.steps, args.warmup) + 1)]

    def step(i, summed=True):
        ws.categorical_set_branch_lengths(q, pi, lengths[i])
        if summed:  # log likelihood summed over the sites on the device
            ws.categorical_log_likelihood(model, thmm.CategoricalInputUpdate.kNone)
        else:  # per-site log Z downloaded
            ws.categorical_logz_forward(model, thmm.CategoricalInputUpdate.kNone)
        return ws.branch_length_backward(q, None)

    def timed(summed=True):
        start = time.perf_counter()
        for i in range(args.steps):
            out = step(i, summed)
        return out, (time.perf_counter() - start) / args.steps * 1e3

    # Warm-up of every variant first, then each timed over the same K steps.
    for leaves_flag in (0, 1):
        ws.set_tuning(input_leaves=leaves_flag)
        for i in range(args.warmup):
            step(i)
            step(i, summed=False)
    grad, ms = timed()
    _, ms_per_site = timed(summed=False)
    ws.set_tuning(input_leaves=0)  # every node factor written first
    _, ms_unfused = timed()
    ws.set_tuning(input_leaves=1)
    start = time.perf_counter()
    for i in range(args.steps):
        ws.categorical_logz_forward(model, thmm.CategoricalInputUpdate.kFactors)
        ws.branch_length_backward(q, None)
    ms_factors = (time.perf_counter() - start) / args.steps * 1e3
    # Parity: the last step's lengths. log Z of 16 sites against the oracle,
    # the FP32 gradient against the FP64 form and against central finite
    # differences of the FP64 form's summed log Z on 4 edges.
    last = lengths[args.steps - 1]
    lz32 = ws.categorical_logz_forward(model, thmm.CategoricalInputUpdate.kNone)
    ws.categorical_set_branch_lengths(q, pi, last)
    lz32 = ws.categorical_logz_forward(model, thmm.CategoricalInputUpdate.kNone)
    grad32 = ws.branch_length_backward(q, None)
    ws64 = thmm.Workspace()
    ws64.reserve_categorical_marginals(plan, k, sites, categories, leaves, precision="f64")
    model64 = thmm.BatchedCategoricalModelView(plan, k, sites, categories, leaves, obs, root, emission,
                                               thmm.expm_reversible(q, pi, last))
    lz64 = ws64.categorical_logz_forward(model64)
    grad64 = ws64.branch_length_backward(q, None)
    fd, picks = [], [0, e // 3, e // 2, e - 1]
    for j in picks:
        h = 1e-5 * max(last[j], 1e-3)
        vals = []
        for sign in (1.0, -1.0):
            tt = last.copy()
            tt[j] += sign * h
            m = thmm.BatchedCategoricalModelView(plan, k, sites, categories, leaves, obs, root, emission,
                                                 thmm.expm_reversible(q, pi, tt))
            vals.append(float(np.sum(ws64.categorical_logz_forward(m))))
        fd.append((vals[0] - vals[1]) / (2 * h))
    fd = np.array(fd)
    picked = np.arange(0, sites, sites // 16)
    nodes = np.ones((picked.size, n, k))
    nodes[:, 0, :] = root.astype(np.float32)
    nodes[:, leaves, :] *= emission[obs[picked]]
    p32 = thmm.expm_reversible(q, pi, last).astype(np.float32).astype(np.float64)
    olz, _, _ = oracle.marginals_batched(plan, k, nodes, p32, threads=max(1, min(CPU_SHARE, host_cpus()["affinity"])))
    parity = {
        "log_z_rel_vs_oracle_16_sites": float(np.max(np.abs(lz32[picked] - olz[:, 0]) / np.maximum(1, np.abs(olz[:, 0])))),
        "log_z_rel_fp32_vs_fp64": float(np.max(np.abs(lz32 - lz64) / np.maximum(1, np.abs(lz64)))),
        "grad_rel_fp32_vs_fp64": float(np.max(np.abs(grad32 - grad64)) / max(1.0, np.max(np.abs(grad64)))),
        "grad_fp64_vs_central_fd_rel": float(np.max(np.abs(grad64[picks] - fd) / np.maximum(1.0, np.abs(fd)))),
        "grad_fp32_vs_central_fd_rel": float(np.max(np.abs(grad32[picks] - fd) / np.maximum(1.0, np.abs(fd)))),
    }
    ws64.close()
    # Roofline of the dominant level kernel: the same levels profiled on the
    # dense form of the model (node factors materialised [B][N][K] FP32).
    dense = torch.ones((sites, n, k), device="cuda")
    dense[:, 0, :] = torch.from_numpy(root.astype(np.float32)).cuda()
    dense[:, torch.from_numpy(leaves.astype(np.int64)).cuda(), :] *= torch.from_numpy(
        emission.astype(np.float32)[obs]).cuda()
    dedges = torch.from_numpy(p32.astype(np.float32)).cuda()
    wsd, outd, calld, dargs = marginal_pass(thmm, torch, plan, k, sites, "f32", dense, dedges)
    calld()
    timings = wsd.profile_marginals_device(*dargs)
    roofline = roofline_from(timings, k, 4, SharedEdgeReads(plan), workload="phylo1024_k4")
    wsd.close()
    log(f"categorical K=4: {ms:.3f} ms/step (per-site log Z {ms_per_site:.3f}, leaves unfused {ms_unfused:.3f}, "
        f"kFactors {ms_factors:.3f}); parity {parity}")
    cpu = None
    if not args.no_cpu:
        sample = sites  # the whole workload: about 0.3 s per pass on 16 threads
        cn = np.ones((sample, n, k))
        cn[:, 0, :] = root
        cn[:, leaves, :] *= emission[obs[:sample]]
        pf = thmm.expm_reversible(q, pi, last)
        ratio = np.matmul(q, pf) / pf
        threads = max(1, min(CPU_SHARE, host_cpus()["affinity"]))

        def chain():
            _, _, em = oracle.marginals_batched(plan, k, cn, pf, threads=threads)
            return np.einsum("bepc,epc->e", em, ratio)

        cpu = cpu_median(f"all {sample} sites ({sample * n} node-sites), FP64 oracle log Z + marginals + "
                         "branch-length chain, one site per thread", sample * n, chain)
    ws.close()
    return result_line(
        "nodes/s log Z + d log Z/d t_e, FP32 categorical (1024 sites x 8191 nodes)", sites * n / (ms * 1e-3), args, ms,
        "f32", "synthetic (SplitMix64, seed 0x54484D4D00000003: observations uniform in {0..3} with 5% missing, root "
        "Dirichlet(1), emission [I;1], reversible Q, t_e ~ Exp(mean 0.1))",
        {"workload": "1024 sites on one balanced binary tree of 4096 leaves (8191 nodes), K=4, FP32 categorical; step = "
         "device expm of the branch lengths + forward (kNone, leaf rakes reading the observations, log likelihood "
         "summed over the sites on the device) + backward to d (sum log Z)/d t_e (host wall clock, synchronous "
         "calls)", "nodes": int(sites * n), "levels": int(plan.primitive_batches.size)},
        roofline=roofline, cpu_baseline=cpu, per_site_log_z_ms_per_step=round(ms_per_site, 4),
        unfused_leaves_ms_per_step=round(ms_unfused, 4), kfactors_ms_per_step=round(ms_factors, 4), parity=parity,
        grad_norm=float(np.linalg.norm(grad)))


def run_dense_state_adjoint(args, thmm, oracle, torch) -> int:
    """Random full binary (Yule) trees of 16,383 nodes, K = 64, log node and
    edge potentials N(0, 1), FP32: log Z plus the gradients of every node and
    edge log potential through the forward/backward pair. A step is one
    forward + backward over the forest."""
    k = 64
    trees = max(1, int(round(256 * args.scale)))
    t0 = time.perf_counter()
    topology = uniforms_host(config_seed(5, STREAM_TOPOLOGY), 0, trees * 8191).reshape(trees, 8191)
    parts = [yule_tree(8192, topology[t]) for t in range(trees)]
    parents, offsets = concatenate(parts)
    plan = thmm.Plan(parents)
    n, e, c = plan.num_nodes, plan.num_edges, plan.num_compressions
    log(f"forest: {trees} Yule trees, {n} nodes, {c} compressions; built and planned in {time.perf_counter() - t0:.1f}s")
    check_device_stream(torch)
    log_nodes = fill_lognormal_device(torch, torch.empty((1, n, k), device="cuda"), config_seed(5, STREAM_NODES), log=True)
    log_edges = fill_lognormal_device(torch, torch.empty((e, k, k), device="cuda"), config_seed(5, STREAM_EDGES), log=True)
    grad_nodes = torch.empty((1, n, k), device="cuda")
    grad_edges = torch.empty((e, k, k), device="cuda")
    log_z = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, 1, precision="f32", device_resident=True)
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        ws.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), 1, log_z.data_ptr(), stream)
        ws.logz_backward_device(0, grad_nodes.data_ptr(), grad_edges.data_ptr(), False, stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        step()
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    value = n / (ms * 1e-3)
    # Algorithmic bytes of one step (values only, s = 4): node staging (read
    # + write), then per rake / compression the forward and reverse figures
    # of op_bytes (kind 0 rakes, 3 compressions). Log edges are read where
    # those figures read an edge: rakes and compressions take them in place.
    s4, rakes = 4, n - c - plan.roots.size
    step_bytes = 2 * n * k * s4 + rakes * (op_bytes(1, 0, k, s4) + op_bytes(3, 0, k, s4)) + c * (
        op_bytes(1, 3, k, s4) + op_bytes(3, 3, k, s4)
    )
    achieved = step_bytes / (ms * 1e-3) / 1e9
    log(f"K=64 FP32 adjoint pair: {ms:.2f} ms/step, {value:.4g} nodes/s, {achieved:.0f} GB/s over the whole step")
    # Parity on two whole trees against the FP64 oracle, and the
    # size-independent normalisation over the forest.
    worst = {"log_z_rel": 0.0, "node_grad_abs": 0.0, "edge_grad_abs": 0.0}
    lz_host = log_z.cpu().numpy()
    for t in (0, trees - 1):
        lo, hi, elo, ehi = tree_slices(offsets, t)
        tree_plan = thmm.Plan(parts[t])
        nn = torch.exp(log_nodes[0, lo:hi].double()).cpu().numpy()
        ee = torch.exp(log_edges[elo:ehi].double()).cpu().numpy()
        olz, onm, oem = oracle.marginals(tree_plan, k, nn, ee)
        worst["log_z_rel"] = max(worst["log_z_rel"], abs(lz_host[t] - olz[0]) / max(1.0, abs(olz[0])))
        worst["node_grad_abs"] = max(worst["node_grad_abs"], float(np.max(np.abs(grad_nodes[0, lo:hi].double().cpu().numpy() - onm))))
        worst["edge_grad_abs"] = max(worst["edge_grad_abs"], float(np.max(np.abs(grad_edges[elo:ehi].double().cpu().numpy() - oem))))
    worst["node_normalisation_abs"] = float((grad_nodes[0].double().sum(1) - 1.0).abs().max())
    worst["trees_checked"] = 2
    log(f"parity: {worst}")
    budgeted = None
    if args.path_budget_frac > 0:
        budgeted = budgeted_pass(args, thmm, torch, plan, k, ws, log_nodes, log_edges, log_z, grad_nodes, grad_edges,
                                 ms)
    ws.close()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_log(thmm, oracle, parts, offsets, log_nodes[0], log_edges, k, trees=min(trees, 16))
    result = {
        "metric": "nodes/s log Z + adjoint (dense-state forest, log potentials)",
        "value": value,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (SplitMix64 / Box-Muller, seed 0x54484D4D00000005: log node and edge potentials N(0,1))",
        "config": {"workload": f"{trees} random full binary (Yule) trees of 16383 nodes, K=64, FP32 log Z + gradients of "
                   "all node and edge log potentials", "nodes": int(n), "compressions": int(c), "scale": args.scale},
        "roofline": {"kernel": "whole step", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_step": int(step_bytes)},
        "cpu_baseline": cpu,
        "parity": worst,
        "path_budget": budgeted,
    }
    print(json.dumps(result), flush=True)
    return 0


def edge_checksum(torch, grad_edges):
    """Sum of the int32 bit patterns of every value (as int64), taken in
    chunks: equal arrays give equal sums."""
    total = 0
    flat = grad_edges.view(grad_edges.shape[0], -1)
    step = max(1, (1 << 27) // flat.shape[1])
    for i in range(0, flat.shape[0], step):
        total += int(flat[i:i + step].view(torch.int32).to(torch.int64).sum())
    return total


def budgeted_pass(args, thmm, torch, plan, k, ws, log_nodes, log_edges, log_z, grad_nodes, grad_edges, ms):
    """The same step with the path storage capped at a fraction of what the
    unbudgeted reservation holds: trees run in groups sharing one path
    store and the backward half recomputes each group's paths. Its outputs
    are compared with the unbudgeted ones (log Z and node gradients bit for
    bit, edge gradients through a checksum of their bit patterns)."""
    _, held_full = ws.path_storage()
    ref_lz, ref_nodes = log_z.clone(), grad_nodes.clone()
    ref_edges = edge_checksum(torch, grad_edges)
    ws.close()
    budget = int(held_full * args.path_budget_frac)
    wb = thmm.Workspace()
    wb.set_path_budget(budget)
    wb.reserve_marginals(plan, k, 1, precision="f32", device_resident=True)
    groups, held = wb.path_storage()
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        wb.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), 1, log_z.data_ptr(), stream)
        wb.logz_backward_device(0, grad_nodes.data_ptr(), grad_edges.data_ptr(), False, stream)

    ms_budget = time_calls(torch, step, args.steps, args.warmup)
    equal = (bool(torch.equal(log_z, ref_lz)) and bool(torch.equal(grad_nodes, ref_nodes))
             and edge_checksum(torch, grad_edges) == ref_edges)
    wb.close()
    out = {"budget_bytes": budget, "groups": int(groups), "path_bytes_held": int(held),
           "path_bytes_unbudgeted": int(held_full), "ms_per_step": round(ms_budget, 4),
           "recompute_cost": round(ms_budget / ms - 1.0, 4), "outputs_equal_unbudgeted": equal}
    log(f"path budget {budget / 1e9:.2f} GB: {groups} tree groups, {held / 1e9:.2f} GB held "
        f"(unbudgeted {held_full / 1e9:.2f} GB), {ms_budget:.2f} ms/step, outputs equal: {equal}")
    return out


def cpu_baseline_log(thmm, oracle, parts, offsets, log_nodes, log_edges, k, trees, repeats=5):
    """The FP64 oracle (log Z + node and edge marginals) on the first `trees`
    whole trees of a forest given log potentials, one tree per task over
    CPU_SHARE host threads; one warm-up pass, then the median of `repeats`."""
    host = host_cpus()
    threads = max(1, min(CPU_SHARE, host["affinity"]))
    jobs = []
    for t in range(trees):
        lo, hi, elo, ehi = tree_slices(offsets, t)
        jobs.append((thmm.Plan(parts[t]), np.exp(log_nodes[lo:hi].double().cpu().numpy()),
                     np.exp(log_edges[elo:ehi].double().cpu().numpy())))
    total = sum(parts[t].size for t in range(trees))

    def work(job):
        plan, nn, ee = job
        return oracle.marginals(plan, k, nn, ee)

    with ThreadPoolExecutor(max_workers=threads) as pool:
        seconds, times = timed_median(lambda: list(pool.map(work, jobs)), repeats)
    return {"value": total / seconds, "unit": "nodes/s", "cores": threads, "kind": "port",
            "sample": f"the first {trees} whole trees ({total} nodes), FP64 oracle log Z + node and edge marginals "
                      f"(the gradients of the log potentials); median of {repeats} passes after one warm-up "
                      f"({seconds:.2f}s per pass)",
            "host": host, "pass_seconds": [round(x, 3) for x in times]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=4, choices=[1, 2, 3, 4, 5],
                    help="4: the uneven K=16 forest (default); 5: the K=64 Yule forest through the adjoint pair")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the trees (quick runs)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-nodes", type=int, default=400_000)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch variants")
    ap.add_argument("--no-half", action="store_true", help="skip the binary16 edge-marginal leg")
    ap.add_argument("--no-map", action="store_true", help="skip the device MAP and sampling legs")
    ap.add_argument("--path-budget-frac", type=float, default=0.25,
                    help="config 5: also time the step with path storage capped at this fraction of the "
                         "unbudgeted reservation's (0: skip)")
    ap.add_argument("--persistent", action="store_true",
                    help="also time the persistent-launch form (switched off in this build: the library refuses it)")
    ap.add_argument("--level-table", default="", metavar="PATH",
                    help="write a markdown table of the profiled pass's launches (config 4)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (the reference point that scales HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only (asked for {args.gpus}); batch items and trees are independent, "
            "so N GPUs would each run their own shard with no exchange",
            file=sys.stderr,
        )
        return 2

    import torch

    import oracle
    import parallel_tree_hmm_amd as thmm

    if not thmm.available():
        print("no HIP device visible", file=sys.stderr)
        return 1
    log(f"device: {thmm.device_description(0)}")
    if args.config == 5:
        return run_dense_state_adjoint(args, thmm, oracle, torch)
    if args.config in (1, 2, 3):
        return {1: run_recursive_fp64, 2: run_long_chains, 3: run_phylogenetic}[args.config](args, thmm, oracle, torch)

    if args.copy_calibration:
        # One plain device-to-device copy of known size: reading FETCH_SIZE and
        # WRITE_SIZE for it gives the counters' bytes per algorithmic byte.
        count = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(count, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {count * 4} bytes read, {count * 4} bytes written")
        # Streaming references on the same buffers (event-timed, best of 5):
        # what plain library kernels reach for read+write and read-only.
        refs = {}
        for name, fn, moved in (("copy", lambda: dst.copy_(src), 8 * count),
                                ("add_in_place", lambda: dst.add_(1.0), 8 * count),
                                ("sum", lambda: src.sum(), 4 * count)):
            best = float("inf")
            for _ in range(5):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                fn()
                b.record()
                torch.cuda.synchronize()
                best = min(best, a.elapsed_time(b))
            refs[name] = round(moved / (best * 1e-3) / 1e9, 1)
        log(f"streaming references (GB/s): {refs}")
        del src, dst
        torch.cuda.empty_cache()

    k = 16
    t0 = time.perf_counter()
    parts = forest_parts(4, scale=args.scale)
    parents, offsets = concatenate(parts)
    plan = thmm.Plan(parents)
    plan_s = time.perf_counter() - t0
    n, e = plan.num_nodes, plan.num_edges
    log(f"forest: {len(parts)} trees, {n} nodes, {plan.num_rounds} rounds, {plan.primitive_batches.size} levels, "
        f"{plan.num_compressions} compressions; built and planned in {plan_s:.1f}s")

    check_device_stream(torch)
    nodes = fill_lognormal_device(torch, torch.empty((n, k), device="cuda"), config_seed(4, STREAM_NODES))
    edges = fill_dirichlet_rows_device(torch, torch.empty((e, k, k), device="cuda"), config_seed(4, STREAM_EDGES))
    torch.cuda.synchronize()

    ws, ms, timings, log_z, node_out, edge_out, tuning = run_gpu(
        thmm, torch, plan, k, "f32", nodes, edges, args.steps, args.warmup, not args.no_tune
    )
    value = n / (ms * 1e-3)
    roofline = roofline_from(timings, k, 4, workload="forest4096_k16" if args.scale == 1.0 else None)
    if args.level_table:
        small, total = level_table(timings, k, 4, args.level_table,
                                   f"Per-launch times: uneven forest, {len(parts)} trees ({n} nodes), K=16, FP32 "
                                   "log Z + adjoint (thmm_profile_marginals_device event pairs)",
                                   ops_of=ws.last_profile_ops)
        log(f"level table: {small:.3f} of {total:.3f} ms in launches of fewer than 256 blocks")
    log(f"autotuned launch variants: {tuning}")
    log(f"FP32 form: {ms:.3f} ms/step, {value:.4g} nodes/s; dominant kernel {roofline['kernel']} "
        f"{roofline['achieved']} GB/s ({roofline['frac']:.1%} of HBM peak)")

    # Parity on a sample of trees (one of each size extreme, a few small).
    log_z_h = log_z.cpu().numpy()
    node_view = node_out.view(n, k)
    edge_view = edge_out.view(e, k, k)
    sizes = np.array([p.size for p in parts])
    picks = [int(np.argmin(sizes))] + [int(t) for t in np.flatnonzero(sizes < 3000)[:6]] + [int(np.argmax(sizes))]
    parity = parity_sample(thmm, oracle, parts, offsets, nodes, edges, k, log_z_h, node_view, edge_view, picks)
    # Size-independent checks over the whole forest: every node marginal sums
    # to one and every edge marginal's row sums equal its parent marginal.
    parity["node_normalisation_abs"] = float((node_view.double().sum(1) - 1.0).abs().max())
    eparents = torch.from_numpy(plan.edge_parents.astype(np.int64)).cuda()
    parity["edge_row_sum_abs"] = float((edge_view.double().sum(2) - node_view.double()[eparents]).abs().max())
    parity["trees_checked"] = len(picks)
    log(f"parity: {parity}")
    log("kernel ms (profiled step): " + json.dumps(roofline["kernel_ms_breakdown"]))
    persistent = None
    if args.persistent:
        # The pass as level launches and as two persistent launches (one per
        # direction) whose blocks pull level chunks from a device queue, with
        # the other tuned launch variants; outputs must be bit-equal.
        kept = ws.tuning()["persistent"]
        stream = torch.cuda.current_stream()
        pargs = (nodes.data_ptr(), edges.data_ptr(), 1, log_z.data_ptr(), node_out.data_ptr(), edge_out.data_ptr(),
                 stream.cuda_stream)
        legs = {}
        for mode in (0, 1):
            ws.set_tuning(persistent=mode)
            legs[mode] = time_calls(torch, lambda: ws.marginals_device(*pargs), args.steps, args.warmup)
            ws.synchronize()
            if mode == 0:
                ref = (log_z.clone(), node_out.clone(), edge_out.clone())
        equal = all(bool(torch.equal(a, b)) for a, b in zip(ref, (log_z, node_out, edge_out)))
        ptimes = ws.profile_marginals_device(*pargs)
        persistent = {"levels_ms_per_step": round(legs[0], 4), "ms_per_step": round(legs[1], 4),
                      "value": n / (legs[1] * 1e-3), "speedup_vs_levels": round(legs[0] / legs[1], 4),
                      "bit_equal_to_levels": equal, "tuned_choice": kept,
                      "launch_ms": {("forward" if ph == 1 else "reverse"): round(t, 4) for ph, pr, c, b, t in ptimes
                                    if pr == 9}}
        log(f"persistent passes: {persistent}")
        ws.set_tuning(persistent=kept)
        del ref
    ws.close()
    del eparents, node_view, edge_view, ws
    torch.cuda.empty_cache()

    half = None
    if not args.no_half:
        # The same pass, with the FP32 form's launch variants, and edge
        # marginals stored as IEEE binary16 (a reservation flag): the
        # arithmetic is the FP32 form's, each edge value is rounded once at
        # the store. Checked against the FP32 output in units of half a
        # binary16 ulp (|h - f| / (2^-11 |f| + 2^-25)).
        ws16, ms16, _, lz16, nout16, eout16, tuning16 = run_gpu(
            thmm, torch, plan, k, "f32", nodes, edges, args.steps, args.warmup, half_edges=True, use=tuning
        )
        ratio, chunk = 0.0, 1 << 26
        for i in range(0, eout16.numel(), chunk):
            h = eout16[i:i + chunk].double()
            f = edge_out[i:i + chunk].double()
            ratio = max(ratio, float(((h - f).abs() / (f.abs() * 2.0**-11 + 2.0**-25)).max()))
            del h, f
        half = {"ms_per_step": round(ms16, 4), "value": n / (ms16 * 1e-3), "speedup_vs_fp32_edges": round(ms / ms16, 3),
                "edge_half_ulps_vs_fp32": round(ratio, 4),
                "log_z_equal": bool(torch.equal(lz16, log_z)), "node_marginals_equal": bool(torch.equal(nout16, node_out)),
                "tuning": tuning16}
        log(f"binary16 edge outputs: {ms16:.3f} ms/step, {ms / ms16:.3f}x; {half}")
        ws16.close()
        del ws16, lz16, nout16, eout16
        torch.cuda.empty_cache()
    del node_out, edge_out, log_z
    torch.cuda.empty_cache()

    decode = None
    if not args.no_map:
        decode = run_map_and_sampling(thmm, torch, plan, k, nodes, edges, args, config_seed(4, STREAM_VARIATES))
        log(f"device MAP and sampling: {decode}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(thmm, oracle, parts, offsets, nodes, edges, k, args.cpu_nodes)
        log(f"CPU oracle: {cpu}")

    fp64 = None
    if not args.no_fp64:
        # The plain FP64 form on the same inputs, widened; the FP32 copies are
        # released first so that both forms fit in HBM one after the other.
        nodes64, edges64 = nodes.double(), edges.double()
        del nodes, edges
        torch.cuda.empty_cache()
        try:
            ws64, ms64, _, lz64, _, _, tuning64 = run_gpu(
                thmm, torch, plan, k, "f64", nodes64, edges64, max(2, args.steps // 2), 1, not args.no_tune
            )
            lz64_h = lz64.cpu().numpy()
            fp64 = {"ms_per_step": round(ms64, 4), "value": n / (ms64 * 1e-3), "fp32_speedup": round(ms64 / ms, 3),
                    "tuning": tuning64,
                    "fp32_vs_fp64_log_z_rel": float(np.max(np.abs(lz64_h - log_z_h) / np.maximum(1, np.abs(lz64_h))))}
            log(f"plain FP64 form: {ms64:.3f} ms/step; FP32 speedup {ms64 / ms:.2f}x")
            ws64.close()
            del ws64, lz64
        except torch.OutOfMemoryError as error:
            fp64 = {"error": f"out of HBM: {error}"[:200]}
        torch.cuda.empty_cache()

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (SplitMix64 / Box-Muller, seed 0x54484D4D00000004: node potentials exp(N(0,1)), "
                "edge rows Dirichlet(1), one variate per node)",
        "config": {
            "workload": "uneven forest: 4096 trees, N log-uniform [16, 65536], random recursive / caterpillar / star, "
            "K=16, FP32 log Z + adjoint (node and edge marginals)",
            "trees": len(parts),
            "nodes": int(n),
            "states": k,
            "levels": int(plan.primitive_batches.size),
            "scale": args.scale,
        },
        "roofline": roofline,
        "tuning": tuning,
        "cpu_baseline": cpu,
        "fp64_plain": fp64,
        "half_edge_outputs": half,
        "persistent": persistent,
        "map_and_sampling": decode,
        "parity": parity,
    }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
