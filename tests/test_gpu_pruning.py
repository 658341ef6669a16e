"""Core-check pruning (GPUSCAN++/GPUScan.cu:303; on-demand decisions
:557-585, :692-719) against the oracle, through the C ABI (needs an MI355X).

Pruning is off by default (sgc_ctx_set_pruning); these tests switch it on.
A pruned pass (one that returns no similarity flags) defers, in the
workgroup tier, every undecided edge whose endpoints are both decided when
its unit starts; after the union-find pass the deferred core-core edges whose
roots differ and the deferred core-non-core edges are decided after all.
Labels and core flags must equal the oracle's (which follows the reference's
own pruned order) whatever was deferred.

Two graphs make deferral and the second round certain: their endpoints are
decided by the degree-only test alone, before any unit starts.
"""
from __future__ import annotations

import random

import numpy as np
import pytest

import bruteforce as bf
import oracle_bind as ob

import efficientstructuralgraphclusteringongpus_amd as sgc

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def pruned_matches_oracle(ctx, g, eps, mu):
    dg = sgc.DeviceGraph(ctx, g)
    try:
        ctx.set_pruning(True)
        r = sgc.scan(ctx, dg, eps, mu)  # no flags requested: pruned when mu >= 1
        full = sgc.scan(ctx, dg, eps, mu, want_sim=True)  # flags requested: every edge decided
    finally:
        ctx.set_pruning(False)
        dg.close()
    labels, core, _, st = ob.orc_scan(g.row_ptr, g.adj, eps, mu)
    assert r.stats.pruned == (1 if mu >= 1 else 0)
    assert full.stats.pruned == 0 and full.stats.deferred_edges == 0 and full.stats.redo_edges == 0
    assert np.array_equal(r.core, core), "core flags differ"
    diff = np.nonzero(r.labels != labels)[0]
    assert len(diff) == 0, f"labels differ at {diff[:10]}"
    for k in ("cores", "clusters", "borders", "hubs", "outliers"):
        assert getattr(r.stats, k) == st[k], k
    return r.stats


def from_edges(n, edges):
    e = np.array(edges, dtype=np.int64)
    return sgc.HostGraph.from_edges(n, e[:, 0].astype(np.int32), e[:, 1].astype(np.int32))


def test_border_needs_deferred_edge(ctx):
    """Vertices 0..9 form a clique (each core by ten degree-only similar
    edges at eps 0.05, mu 8); vertices 0..6 are also joined to hub 10, whose
    other 1999 neighbours have degree 1, so the degree-only test leaves the
    hub fewer than mu possibly similar edges (not core).  Edge (0, 10) is
    deferred (both ends decided), then needed for the border rule: it is
    similar (6 common neighbours: 8^2 = 64 >= 0.05^2 * 11 * 2007), so the hub
    joins cluster 0 as a border."""
    edges = [(i, j) for i in range(10) for j in range(i + 1, 10)]
    edges += [(i, 10) for i in range(7)]
    edges += [(10, 11 + t) for t in range(1999)]
    g = from_edges(2010, edges)
    st = pruned_matches_oracle(ctx, g, 0.05, 8)
    say("border case", st.deferred_edges, st.redo_edges, st.borders)
    assert st.redo_edges >= 1 and st.borders >= 1


def test_union_needs_deferred_edge(ctx):
    """Hubs 0 and 1 (degree 1101) are joined to each other, to 20 shared
    degree-2 vertices and to 1080 leaves each; at eps 0.01 every hub-leaf edge
    is similar by degree alone, so both hubs are cores before any unit runs
    (mu 3) and edge (0, 1) is deferred.  No other similar core-core edge joins
    them, so the union-find pass needs it: 20 common neighbours make it
    similar (22^2 = 484 >= 0.01^2 * 1102^2) and the two hubs form one
    cluster."""
    edges = [(0, 1)]
    shared = list(range(2, 22))
    edges += [(h, s) for h in (0, 1) for s in shared]
    nxt = 22
    for h in (0, 1):
        edges += [(h, nxt + t) for t in range(1080)]
        nxt += 1080
    g = from_edges(nxt, edges)
    st = pruned_matches_oracle(ctx, g, 0.01, 3)
    say("union case", st.deferred_edges, st.redo_edges, st.clusters)
    assert st.redo_edges >= 1 and st.clusters == 1


def test_pruned_random_graphs(ctx):
    rng = random.Random(33)
    for trial in range(40):
        n = rng.randint(2, 300)
        p = rng.choice([0.02, 0.1, 0.3])
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
        eps = rng.choice([0.1, 0.3, 0.5, 0.7])
        mu = rng.choice([0, 1, 2, 3, 5])
        say("random", trial, n, len(edges), eps, mu)
        rp, adj = bf.csr(n, edges)
        pruned_matches_oracle(ctx, sgc.HostGraph.from_csr(rp, adj), eps, mu)


@pytest.mark.parametrize("scale,eps,mu", [(14, 0.2, 2), (14, 0.2, 5), (16, 0.2, 2), (16, 0.3, 10), (16, 0.6, 3)])
def test_pruned_rmat(ctx, scale, eps, mu):
    g = sgc.HostGraph.rmat(scale, 16, seed=scale)
    st = pruned_matches_oracle(ctx, g, eps, mu)
    say("rmat", scale, eps, mu, "deferred", st.deferred_edges, "redo", st.redo_edges)


def test_pruning_switch(ctx):
    g = sgc.HostGraph.rmat(13, 16, seed=2)
    dg = sgc.DeviceGraph(ctx, g)
    try:
        off = sgc.scan(ctx, dg, 0.3, 3)  # the default
        ctx.set_pruning(True)
        on = sgc.scan(ctx, dg, 0.3, 3)
    finally:
        ctx.set_pruning(False)
        dg.close()
    assert on.stats.pruned == 1 and off.stats.pruned == 0 and off.stats.deferred_edges == 0
    assert np.array_equal(on.labels, off.labels) and np.array_equal(on.core, off.core)
