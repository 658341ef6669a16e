"""HIP path vs the oracle, through the C ABI (needs an MI355X).

Bit-exact: per-vertex labels (cluster id = minimum core id, canonical
minimum-candidate borders, -1 hub, -2 outlier), core flags and the
similarity flag of every directed entry.
"""
from __future__ import annotations

import json
import os
import random
import subprocess
import sys

import numpy as np
import pytest

import bruteforce as bf
import oracle_bind as ob
from label_check import labels_from_flags

import efficientstructuralgraphclusteringongpus_amd as sgc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print("[case]", *a, flush=True)


def run_gpu(ctx, g, eps, mu):
    dg = sgc.DeviceGraph(ctx, g)
    try:
        return sgc.scan(ctx, dg, eps, mu, want_sim=True)
    finally:
        dg.close()


def assert_matches_oracle(ctx, g, eps, mu):
    rp, adj = np.array(g.row_ptr), np.array(g.adj)
    labels, core, sim, st = ob.orc_scan(rp, adj, eps, mu)
    r = run_gpu(ctx, g, eps, mu)
    assert np.array_equal(r.sim, sim), f"similarity differs at {np.nonzero(r.sim != sim)[0][:10]}"
    assert np.array_equal(r.core, core)
    assert np.array_equal(r.labels, labels), f"labels differ at {np.nonzero(r.labels != labels)[0][:10]}"
    assert r.stats.cores == st["cores"]
    assert r.stats.clusters == st["clusters"]
    assert r.stats.similar_edges == st["similar_edges"]
    assert r.stats.hubs == st["hubs"] and r.stats.outliers == st["outliers"] and r.stats.borders == st["borders"]
    assert r.stats.u_edges == st["prep_u"]
    return r


def test_known_answer_graphs(ctx):
    data = json.load(open(os.path.join(GOLDEN, "kat_graphs.json")))
    for case in data["cases"]:
        say(case["name"])
        rp, adj = bf.csr(case["n"], [tuple(e) for e in case["edges"]])
        g = sgc.HostGraph.from_csr(rp, adj)
        r = run_gpu(ctx, g, case["eps"], case["mu"])
        assert list(r.labels) == case["labels"], case["name"]
        assert list(r.core) == case["core"], case["name"]
        assert r.stats.cores == case["core_number"] and r.stats.clusters == case["cluster_number"]


def test_random_small_graphs(ctx):
    rng = random.Random(21)
    for trial in range(60):
        n = rng.randint(1, 300)
        p = rng.choice([0.01, 0.05, 0.2, 0.5])
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
        eps = rng.choice([0.0, 0.1, 0.25, 0.5, 0.65, 0.9, 1.0, 1.5])
        mu = rng.choice([-2, 0, 1, 2, 3, 6])
        say("random", trial, n, len(edges), eps, mu)
        rp, adj = bf.csr(n, edges)
        assert_matches_oracle(ctx, sgc.HostGraph.from_csr(rp, adj), eps, mu)


def test_rmat16_default_config(ctx):
    say("rmat16 eps=0.5 mu=5")
    g = sgc.HostGraph.rmat(16, 16, seed=1)
    r = assert_matches_oracle(ctx, g, 0.5, 5)
    # R-MAT at eps = 0.5 is almost structureless: 350 similar edges, no core
    assert r.stats.similar_edges == 350 and r.stats.cores == 0
    say("rmat16 eps=0.2 mu=2")
    r = assert_matches_oracle(ctx, g, 0.2, 2)
    assert r.stats.cores > 0 and r.stats.clusters > 0 and r.stats.borders > 0 and r.stats.hubs > 0


@pytest.mark.parametrize("eps", [0.2, 0.4, 0.6, 0.8])
@pytest.mark.parametrize("mu", [2, 5, 10])
def test_rmat14_grid(ctx, eps, mu):
    say("rmat14", eps, mu)
    g = sgc.HostGraph.rmat(14, 16, seed=2)
    assert_matches_oracle(ctx, g, eps, mu)


@pytest.mark.parametrize("scale,ef,eps", [(16, 16, 0.2), (18, 16, 0.5), (17, 32, 0.3)])
def test_tiered_and_wave_kernels_agree(ctx, scale, ef, eps):
    """Two independent similarity kernels (LDS-hash tiers; one wave per edge
    with binary search) give identical flags; a sample is checked against the
    oracle too."""
    say("kernels agree", scale, ef, eps)
    g = sgc.HostGraph.rmat(scale, ef, seed=3)
    dg = sgc.DeviceGraph(ctx, g)
    try:
        ctx.set_similarity_kernel(sgc.Context.SIM_WAVE_PER_EDGE)
        a = sgc.scan(ctx, dg, eps, 3, want_sim=True)
        ctx.set_similarity_kernel(sgc.Context.SIM_TIERED)
        b = sgc.scan(ctx, dg, eps, 3, want_sim=True)
    finally:
        ctx.set_similarity_kernel(sgc.Context.SIM_TIERED)
        dg.close()
    assert np.array_equal(a.sim, b.sim)
    assert np.array_equal(a.labels, b.labels)
    rp, adj = np.array(g.row_ptr), np.array(g.adj)
    sample = np.random.default_rng(1).choice(len(adj), size=min(len(adj), 200_000), replace=False)
    _, ref = ob.orc_similarity_entries(rp, adj, eps, sample.astype(np.uint64))
    assert np.array_equal(b.sim[sample], ref)


def _kernels_agree(ctx, g, eps, mu, sample_size=200_000):
    dg = sgc.DeviceGraph(ctx, g)
    try:
        ctx.set_similarity_kernel(sgc.Context.SIM_WAVE_PER_EDGE)
        a = sgc.scan(ctx, dg, eps, mu, want_sim=True)
        ctx.set_similarity_kernel(sgc.Context.SIM_TIERED)
        b = sgc.scan(ctx, dg, eps, mu, want_sim=True)
    finally:
        ctx.set_similarity_kernel(sgc.Context.SIM_TIERED)
        dg.close()
    assert np.array_equal(a.sim, b.sim)
    assert np.array_equal(a.labels, b.labels)
    rp, adj = np.array(g.row_ptr), np.array(g.adj)
    sample = np.random.default_rng(5).choice(len(adj), size=min(len(adj), sample_size), replace=False)
    _, ref = ob.orc_similarity_entries(rp, adj, eps, sample.astype(np.uint64))
    assert np.array_equal(b.sim[sample], ref)
    return b


@pytest.mark.parametrize("eps", [0.2, 0.6])
def test_windows_beyond_one_bitmap(ctx, eps):
    """n > 2^20: owner rows of the large tier span two 2^20-id bitmap windows;
    each short row's slice per window is found while streaming."""
    say("rmat21 windows", eps)
    g = sgc.HostGraph.rmat(21, 16, seed=7)
    assert g.n > (1 << 20) and g.max_degree > 4 * 8192
    _kernels_agree(ctx, g, eps, 3)


def test_bitmap_windows_for_very_long_owner_rows(ctx):
    """n > 2^20 and owner rows of 140K keys spread over both 2^20-id windows,
    with the mirror entries in either window."""
    say("bitmap windows")
    n = 1_100_000
    k = 140_000
    rng = np.random.default_rng(9)
    common = np.arange(2, 2 + k, dtype=np.int64) * 7 % (n - 2) + 2  # spread over both windows
    common = np.unique(common)
    bg_src = rng.integers(0, n, size=2_000_000)
    bg_dst = rng.integers(0, n, size=2_000_000)
    src = np.concatenate([[0], np.zeros(len(common)), np.ones(len(common)), bg_src]).astype(np.int32)
    dst = np.concatenate([[1], common, common, bg_dst]).astype(np.int32)
    g = sgc.HostGraph.from_edges(n, src, dst)
    assert g.max_degree > 131_072
    for eps in (0.3, 0.9):
        r = _kernels_agree(ctx, g, eps, 2, sample_size=400_000)
    labels, core, sim, st = ob.orc_scan(np.array(g.row_ptr), np.array(g.adj), 0.9, 2)
    assert np.array_equal(r.sim, sim) and np.array_equal(r.labels, labels)


def test_window_runs_in_one_hash(ctx):
    """n > 3 * 2^20 (four windows): owner rows over 8192 keys are cut into runs
    of consecutive windows holding at most 8192 keys (one cuckoo hash per
    run) and single windows beyond that (bitmap); the owners, and so the
    mirror entries, sit in different windows."""
    say("window runs")
    n, W = 3_300_000, 1 << 20
    rng = np.random.default_rng(11)
    hubs = np.array([1_500_000, 2_500_001, 3_100_000])

    def spread(counts):
        parts = [rng.choice(np.arange(w * W, min(n, (w + 1) * W)), size=c, replace=False) for w, c in enumerate(counts)]
        return np.setdiff1d(np.concatenate(parts), hubs)

    k0 = spread([2000, 3000, 3000, 4000])  # run of windows 0-2, then window 3 alone
    k1 = np.union1d(k0[rng.random(len(k0)) < 0.8], spread([500, 500, 500, 500]))
    k2 = spread([1000, 1000, 9000, 1000])  # run of windows 0-1, window 2 (bitmap), window 3
    mids = np.setdiff1d(rng.choice(n, size=64, replace=False), hubs)
    pool = np.union1d(k0, k2)
    src, dst = [], []
    for h, keys in zip(hubs, (k0, k1, k2)):
        src.append(np.full(len(keys), h))
        dst.append(keys)
    for x in mids:
        nb = rng.choice(pool, size=int(rng.integers(1500, 3000)), replace=False)
        src.append(np.full(len(nb) + 3, x))
        dst.append(np.concatenate([nb, hubs]))
    src.append(np.array([hubs[0], hubs[1]]))
    dst.append(np.array([hubs[1], hubs[2]]))
    src.append(rng.integers(0, n, size=1_000_000))
    dst.append(rng.integers(0, n, size=1_000_000))
    g = sgc.HostGraph.from_edges(n, np.concatenate(src).astype(np.int32), np.concatenate(dst).astype(np.int32))
    assert g.n > 3 * W and g.max_degree > 8192
    for eps in (0.3, 0.6):
        r = _kernels_agree(ctx, g, eps, 2, sample_size=300_000)
    labels, core, sim, st = ob.orc_scan(np.array(g.row_ptr), np.array(g.adj), 0.6, 2)
    assert np.array_equal(r.sim, sim) and np.array_equal(r.labels, labels)


def test_rmat22_four_windows(ctx):
    """R-MAT scale 22 (four 2^20-id windows): both similarity kernels agree
    and a sample matches the oracle."""
    say("rmat22 windows")
    g = sgc.HostGraph.rmat(22, 16, seed=4)
    assert g.n > 3 * (1 << 20)
    _kernels_agree(ctx, g, 0.3, 4, sample_size=100_000)


def test_planted_partition_hubs(ctx):
    """Hub-heavy graph: 64-vertex-scale version of the planted-partition
    workload (hub rows take the segmented workgroup tier)."""
    say("planted")
    g = sgc.HostGraph.planted(n=200_000, communities=2_000, hubs=16, hub_fraction=0.1, seed=2)
    rp, adj = np.array(g.row_ptr), np.array(g.adj)
    r = run_gpu(ctx, g, 0.5, 4)
    sample = np.random.default_rng(2).choice(len(adj), size=300_000, replace=False).astype(np.uint64)
    _, ref = ob.orc_similarity_entries(rp, adj, 0.5, sample)
    assert np.array_equal(r.sim[sample.astype(np.int64)], ref)
    labels, core = labels_from_flags(rp, adj, r.sim, 4)
    assert np.array_equal(labels, r.labels) and np.array_equal(core, r.core)


def test_int32_wrap_hub_pair(ctx):
    """Two hubs with 50000 common neighbours: (cn+2)^2 wraps in the reference's
    int product and the hub-hub edge comes out dissimilar."""
    say("wrap hub pair")
    k = 50000
    src = [0] + [0] * k + [1] * k + list(range(2, 2 + k - 1))
    dst = [1] + list(range(2, 2 + k)) + list(range(2, 2 + k)) + list(range(3, 3 + k - 1))
    g = sgc.HostGraph.from_edges(k + 2, src, dst)
    for eps in (0.2, 0.5, 0.9):
        r = assert_matches_oracle(ctx, g, eps, 2)
        assert r.sim[0] == 0  # entry (0 -> 1)


def test_long_paths_and_cycles_are_one_cluster(ctx):
    """Regression: deep union-find trees while many threads compress paths;
    every core must end with its component's minimum id."""
    say("long path / cycle")
    for n in (1000, 100_000):
        src = list(range(n - 1)) + [n - 1]
        dst = list(range(1, n)) + [0]
        g = sgc.HostGraph.from_edges(n, src, dst)
        r = assert_matches_oracle(ctx, g, 0.4, 2)
        assert r.stats.clusters == 1 and set(r.labels.tolist()) == {0}


def test_empty_and_isolated(ctx):
    say("edgeless")
    g = sgc.HostGraph.from_csr([0, 0, 0, 0], [])
    for mu in (0, 1, 3):
        r = run_gpu(ctx, g, 0.5, mu)
        assert list(r.labels) == ([-1] * 3 if mu <= 0 else [-2] * 3)


def test_repeat_runs_identical(ctx):
    say("determinism")
    g = sgc.HostGraph.rmat(15, 16, seed=5)
    dg = sgc.DeviceGraph(ctx, g)
    a = sgc.scan(ctx, dg, 0.4, 3, want_sim=True)
    b = sgc.scan(ctx, dg, 0.4, 3, want_sim=True)
    dg.close()
    assert np.array_equal(a.labels, b.labels) and np.array_equal(a.sim, b.sim)


def test_pageable_and_page_locked_uploads_agree(ctx):
    """The upload stages a pageable graph through the context's page-locked
    buffers in 64 MB chunks (R-MAT-20 at edge factor 16 takes two) and sends
    a page-locked one as it lies: the same results either way, in a second
    context too (its own staging), and the same as the oracle on a sample."""
    say("pageable vs page-locked upload")
    g = sgc.HostGraph.rmat(20, 16, seed=3)
    assert g.m * 4 > 64 << 20
    a = run_gpu(ctx, g, 0.4, 4)
    with sgc.Context(0) as ctx2:
        b = run_gpu(ctx2, g, 0.4, 4)
    g.pin()
    c = run_gpu(ctx, g, 0.4, 4)
    for r in (b, c):
        assert np.array_equal(a.sim, r.sim) and np.array_equal(a.labels, r.labels)
    rp, adj = np.array(g.row_ptr), np.array(g.adj)
    sample = np.random.default_rng(3).choice(len(adj), size=200_000, replace=False).astype(np.uint64)
    _, ref = ob.orc_similarity_entries(rp, adj, 0.4, sample)
    assert np.array_equal(a.sim[sample.astype(np.int64)], ref)


def test_rmat20_sampled_oracle_and_label_rules(ctx):
    """Full-size consistency: similarity of a random 2% of entries against the
    oracle, and labels re-derived from the GPU's own flags by the vectorised
    label rules."""
    say("rmat20")
    g = sgc.HostGraph.rmat(20, 16, seed=1)
    rp, adj = np.array(g.row_ptr), np.array(g.adj)
    r = run_gpu(ctx, g, 0.5, 5)
    rng = np.random.default_rng(0)
    sample = rng.choice(len(adj), size=len(adj) // 50, replace=False).astype(np.uint64)
    _, ref = ob.orc_similarity_entries(rp, adj, 0.5, sample)
    assert np.array_equal(r.sim[sample.astype(np.int64)], ref)
    labels, core = labels_from_flags(rp, adj, r.sim, 5)
    assert np.array_equal(core, r.core)
    assert np.array_equal(labels, r.labels)


def test_reference_entry_point(ctx):
    """sgc_GPUScan: the reference's int* GPUScan(graph_t*, ...) contract."""
    import ctypes as C

    say("sgc_GPUScan")
    g = sgc.HostGraph.rmat(12, 16, seed=8)
    lib = sgc.load_library()
    rp32 = np.array(g.row_ptr, dtype=np.uint32)
    adj = np.array(g.adj, dtype=np.int32)
    gt = sgc._GraphT(g.n, g.m, adj.ctypes.data_as(C.POINTER(C.c_int32)), rp32.ctypes.data_as(C.POINTER(C.c_uint32)),
                     None)
    out = lib.sgc_GPUScan(C.byref(gt), None, None, 0.5, 4)
    assert out
    got = np.ctypeslib.as_array(out, shape=(g.n,)).copy()
    C.CDLL(None).free(C.cast(out, C.c_void_p))
    labels, _, _, _ = ob.orc_scan(g.row_ptr, adj, 0.5, 4)
    assert np.array_equal(got, labels)


def test_context_memory(ctx):
    free, total = ctx.memory()
    assert 0 < free <= total and total >= 1 << 30


def test_cli_matches_reference_lines(tmp_path):
    say("GPUScan CLI")
    g = sgc.HostGraph.rmat(13, 16, seed=6)
    folder = str(tmp_path / "g")
    g.save(folder)
    out = str(tmp_path / "labels.bin")
    cli = os.path.join(os.path.dirname(sgc.library_path()), "GPUScan")
    proc = subprocess.run([cli, folder, "0.5", "3", "--out", out], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stderr
    labels, core, _, st = ob.orc_scan(g.row_ptr, g.adj, 0.5, 3)
    assert f"core number: {st['cores']}" in proc.stderr
    assert f"Cluster Number: {st['clusters']}" in proc.stderr
    assert f"max degree: {g.max_degree}" in proc.stdout
    raw = open(out, "rb").read()
    assert np.array_equal(np.frombuffer(raw[: 4 * g.n], dtype=np.int32), labels)
    assert np.array_equal(np.frombuffer(raw[4 * g.n:], dtype=np.uint8), core)
    res_dir = str(tmp_path / "res")
    proc = subprocess.run([cli, folder, "0.5", "3", "--out-dir", res_dir, "--text"], capture_output=True, text=True,
                          timeout=120)
    assert proc.returncode == 0, proc.stderr
    back = sgc.read_result(res_dir)
    assert np.array_equal(back["labels"], labels) and np.array_equal(back["core"], core)
    assert back["summary"]["Cluster Number"] == str(st["clusters"])
    bad = subprocess.run([cli, folder], capture_output=True, text=True)
    assert bad.returncode == 1 and "parameter eps and miu" in bad.stderr
    # out of core within a small budget (the reference's GPUSCAN++O command
    # line takes the same arguments): the same labels and count lines
    budget = max(sgc.ooc_min_budget(g, 0.5) * 1.05, sgc.ooc_min_budget(g, 0.5) + g.m * 11)
    ooc_out = str(tmp_path / "labels_ooc.bin")
    proc = subprocess.run([cli, folder, "0.5", "3", "--out", ooc_out, "--device-budget", str(budget / 2**30)],
                          capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stderr
    assert "out of core within" in proc.stderr and "similarity chips:" in proc.stderr
    assert f"core number: {st['cores']}" in proc.stderr and f"Cluster Number: {st['clusters']}" in proc.stderr
    assert open(ooc_out, "rb").read() == raw
