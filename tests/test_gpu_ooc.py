"""Out-of-core tier: bit-identical to the resident path under device budgets
that force many chips (needs an MI355X)."""
from __future__ import annotations

import json
import os
import random

import numpy as np
import pytest

import bruteforce as bf
import oracle_bind as ob

import efficientstructuralgraphclusteringongpus_amd as sgc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print("[case]", *a, flush=True)


def resident(ctx, g, eps, mu):
    dg = sgc.DeviceGraph(ctx, g)
    try:
        return sgc.scan(ctx, dg, eps, mu, want_sim=True)
    finally:
        dg.close()


def budget_for(g, eps, chip_entries):
    """The smallest accepted budget plus 44 bytes per `chip_entries`: about
    that many more entries per block for chips of two single blocks (43
    bytes per block entry), fewer per block once blocks are grouped."""
    return sgc.ooc_min_budget(g, eps) + 44 * int(chip_entries)


def assert_same(a, b):
    assert np.array_equal(a.sim, b.sim), f"similarity differs at {np.nonzero(a.sim != b.sim)[0][:10]}"
    assert np.array_equal(a.core, b.core)
    assert np.array_equal(a.labels, b.labels), f"labels differ at {np.nonzero(a.labels != b.labels)[0][:10]}"
    for f in ("u_edges", "similar_edges", "cores", "clusters", "borders", "hubs", "outliers"):
        assert getattr(a.stats, f) == getattr(b.stats, f), f


@pytest.mark.parametrize("scale,ef,eps,mu,parts", [(14, 16, 0.3, 3, 6), (16, 16, 0.5, 5, 8), (16, 8, 0.2, 2, 5),
                                                   (17, 16, 0.7, 2, 12)])
def test_ooc_matches_resident_rmat(ctx, scale, ef, eps, mu, parts):
    say("ooc rmat", scale, ef, eps, mu, parts)
    g = sgc.HostGraph.rmat(scale, ef, seed=4)
    ref = resident(ctx, g, eps, mu)
    budget = budget_for(g, eps, g.m // parts)
    r, info = sgc.scan_ooc(ctx, g, eps, mu, budget, want_sim=True)
    say("chips", info.similarity_chips, info.row_chips)
    assert info.similarity_chips > 1
    assert info.device_bytes <= budget
    assert_same(r, ref)


def test_ooc_many_windows_matches_resident(ctx):
    """n > 2^20, so neighbour ids span four 2^20-id windows and the
    workgroup tier makes one pass per window an owner row touches.  Chips
    have no window-offset index and find where each short row's slice of a
    pass ends while streaming it; the resident pass reads the slices from
    the index.  Flags and labels must be identical."""
    g = sgc.HostGraph.rmat(22, 4, seed=9)
    say("many windows", g.n, g.m, g.max_degree)
    assert g.n > 3 << 20 and g.max_degree > 1024
    for eps, mu in ((0.2, 3), (0.5, 2)):
        ref = resident(ctx, g, eps, mu)
        budget = budget_for(g, eps, g.m // 6)
        r, info = sgc.scan_ooc(ctx, g, eps, mu, budget, want_sim=True)
        say("chips", info.similarity_chips, eps, mu)
        assert info.similarity_chips > 1
        assert_same(r, ref)


def test_ooc_from_a_pinned_graph(ctx):
    """Row blocks sent straight from a page-locked graph give the results of
    the staged path, over many similarity chips and row chips."""
    say("ooc pinned")
    g = sgc.HostGraph.rmat(16, 16, seed=7)
    digest = g.digest()
    ref = resident(ctx, g, 0.4, 3)
    budget = budget_for(g, 0.4, g.m // 10)
    staged, info0 = sgc.scan_ooc(ctx, g, 0.4, 3, budget, want_sim=True)
    g.pin()
    assert g.pinned and g.digest() == digest
    g.pin()  # a second call does nothing
    pinned, info1 = sgc.scan_ooc(ctx, g, 0.4, 3, budget, want_sim=True)
    say("chips", info1.similarity_chips, info1.row_chips)
    assert info1.similarity_chips == info0.similarity_chips > 1
    assert_same(staged, ref)
    assert_same(pinned, ref)


@pytest.mark.parametrize("eps", [0.8, 0.95])
def test_ooc_row_chips(ctx, eps):
    """Few undecided edges keep the minimum budget small, so the passes after
    similarity also run over several row chips."""
    say("ooc row chips", eps)
    g = sgc.HostGraph.rmat(16, 16, seed=6)
    ref = resident(ctx, g, eps, 2)
    r, info = sgc.scan_ooc(ctx, g, eps, 2, budget_for(g, eps, g.m // 40), want_sim=True)
    say("chips", info.similarity_chips, info.row_chips)
    assert info.similarity_chips > 1 and info.row_chips > 1
    assert_same(r, ref)


def test_ooc_one_chip_when_the_budget_allows(ctx):
    say("ooc one chip")
    g = sgc.HostGraph.rmat(15, 16, seed=5)
    ref = resident(ctx, g, 0.4, 3)
    r, info = sgc.scan_ooc(ctx, g, 0.4, 3, 1 << 33, want_sim=True)
    assert info.similarity_chips == 1 and info.row_chips == 1
    assert_same(r, ref)
    assert r.stats.ms_core > 0  # per-phase device times are reported for one chip


def test_ooc_known_answers_and_small_graphs(ctx):
    data = json.load(open(os.path.join(GOLDEN, "kat_graphs.json")))
    for case in data["cases"]:
        say("ooc kat", case["name"])
        rp, adj = bf.csr(case["n"], [tuple(e) for e in case["edges"]])
        g = sgc.HostGraph.from_csr(rp, adj)
        r, _ = sgc.scan_ooc(ctx, g, case["eps"], case["mu"], budget_for(g, case["eps"], 4))
        assert list(r.labels) == case["labels"], case["name"]
        assert list(r.core) == case["core"], case["name"]
    rng = random.Random(8)
    for trial in range(25):
        n = rng.randint(1, 200)
        p = rng.choice([0.02, 0.1, 0.3])
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
        eps = rng.choice([0.1, 0.3, 0.5, 0.8])
        mu = rng.choice([0, 1, 2, 4])
        say("ooc random", trial, n, len(edges), eps, mu)
        rp, adj = bf.csr(n, edges)
        g = sgc.HostGraph.from_csr(rp, adj)
        labels, core, sim, st = ob.orc_scan(np.array(rp), np.array(adj), eps, mu)
        r, info = sgc.scan_ooc(ctx, g, eps, mu, budget_for(g, eps, g.m // 4), want_sim=True)
        assert g.m < 400 or info.similarity_chips > 1
        assert np.array_equal(r.labels, labels) and np.array_equal(r.core, core) and np.array_equal(r.sim, sim)


def test_ooc_hub_heavy_planted(ctx):
    say("ooc planted")
    g = sgc.HostGraph.planted(n=100_000, communities=1_000, hubs=8, hub_fraction=0.1, seed=3)
    ref = resident(ctx, g, 0.5, 4)
    r, info = sgc.scan_ooc(ctx, g, 0.5, 4, budget_for(g, 0.5, g.m // 4), want_sim=True)
    assert info.similarity_chips > 1
    assert_same(r, ref)


def group_chips(blocks, k):
    """Similarity chips of `blocks` row blocks in groups of k: each group
    alone, then with every later block."""
    return sum(1 + blocks - min(first + k, blocks) for first in range(0, blocks, k))


def test_ooc_groups_of_blocks(ctx):
    """Where the budget allows, four consecutive row blocks form a group that
    stays on the device (flags included) while every later block passes
    through: the group's earlier chips' flags must survive its later chips."""
    g = sgc.HostGraph.rmat(16, 16, seed=9)
    for eps, mu in ((0.3, 3), (0.6, 2)):
        ref = resident(ctx, g, eps, mu)
        r, info = sgc.scan_ooc(ctx, g, eps, mu, budget_for(g, eps, g.m // 3), want_sim=True)
        say("groups", eps, info.row_blocks, info.blocks_per_group, info.similarity_chips)
        assert info.blocks_per_group == 4 and info.row_blocks > 4
        assert info.similarity_chips == group_chips(info.row_blocks, 4)
        assert_same(r, ref)


def test_ooc_single_blocks_at_the_minimum_budget(ctx):
    """At the smallest budget a chip holds two blocks of the longest row, so
    blocks are not grouped: every pair of blocks is one chip."""
    g = sgc.HostGraph.rmat(13, 8, seed=10)
    r, info = sgc.scan_ooc(ctx, g, 0.4, 2, sgc.ooc_min_budget(g, 0.4), want_sim=True)
    say("single blocks", info.row_blocks, info.blocks_per_group, info.similarity_chips)
    assert info.blocks_per_group == 1 and info.row_blocks > 2
    assert info.similarity_chips == group_chips(info.row_blocks, 1)
    assert_same(r, resident(ctx, g, 0.4, 2))


def test_ooc_rejects_budgets_below_the_minimum(ctx):
    g = sgc.HostGraph.rmat(12, 8, seed=1)
    for budget in (1024, sgc.ooc_min_budget(g, 0.5) - 64 * 1024):
        with pytest.raises(sgc.SGCError) as e:
            sgc.scan_ooc(ctx, g, 0.5, 2, budget)
        assert sgc.SGCError.CODES[e.value.code] == "ENOMEM"
    r, _ = sgc.scan_ooc(ctx, g, 0.5, 2, sgc.ooc_min_budget(g, 0.5))
    assert np.array_equal(r.labels, resident(ctx, g, 0.5, 2).labels)


def test_ooc_edgeless_graph(ctx):
    g = sgc.HostGraph.from_csr(np.zeros(11, dtype=np.uint64), np.zeros(0, dtype=np.int32))
    r, _ = sgc.scan_ooc(ctx, g, 0.5, 2, 1 << 20)
    assert list(r.labels) == [-2] * 10
