"""HIP path at the configured full-size workloads (needs an MI355X).

At these sizes the sequential oracle cannot run a whole pass, so each pass is
checked in two independent parts:

- similarity flags: every entry whose edge takes the int32-wrap branch of
  the reference's predicate (short row longer than 46338, GPUSCAN++/
  GPUScan.cu:220-221) plus a uniform 0.5 % sample of all entries, against
  the oracle's per-edge similarity (prep + merge count + predicate);
- the same pass without similarity flags (which runs the reference's pruned
  core check, GPUScan.cu:303) gives the same labels and core flags;
- labels and core flags: the oracle's phases after similarity
  (orc_labels_from_sim: core check, union-find with minimum-id roots,
  canonical borders, hubs/outliers, GPUScan.cu:276-798) run over the device's
  own full flag array, against the device's labels and core flags; the
  degree-only pre-decision counts against the oracle's.

R-MAT-24 (edge factor 16, seed 1) over the whole eps x mu grid of the
benchmark; the planted-partition graph (n = 2M, 20,000 Zipf(1.5)
communities, 64 hubs x 5 %) at its configured eps 0.5, mu 4 and at eps 0.08
and 0.05, where similar edges exist (at 0.5 none do).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_bind as ob

import efficientstructuralgraphclusteringongpus_amd as sgc

pytestmark = pytest.mark.gpu

WRAP_ROW = 46338  # (s + 2)^2 >= 2^31 beyond this short-row length


def say(*a):
    print("[case]", *a, flush=True)


def wrap_entries(rp, adj):
    deg = np.diff(rp)
    src = np.repeat(np.arange(len(deg), dtype=np.uint32), deg.astype(np.int64))
    short = np.minimum(deg[src], deg[adj])
    return np.nonzero(short > WRAP_ROW)[0].astype(np.uint64)


def check_grid(ctx, g, grid, seed):
    rp, adj = np.asarray(g.row_ptr), np.asarray(g.adj)
    wrap = wrap_entries(rp, adj)
    rng = np.random.default_rng(seed)
    dg = sgc.DeviceGraph(ctx, g)
    out = {}
    try:
        for eps, mus in grid:
            sample = np.union1d(wrap, rng.integers(0, g.m, size=g.m // 200).astype(np.uint64))
            _, ref = ob.orc_similarity_entries(rp, adj, eps, sample)
            pn, py, pu = ob.orc_prep_counts(rp, adj, eps)
            for mu in mus:
                say(f"n={g.n} m={g.m} eps={eps} mu={mu}: {len(sample)} sampled entries, {len(wrap)} on the wrap path")
                r = sgc.scan(ctx, dg, eps, mu, want_sim=True)
                bad = np.nonzero(r.sim[sample.astype(np.int64)] != ref)[0]
                assert len(bad) == 0, f"similarity differs at entries {sample[bad[:10]]}"
                assert r.stats.u_edges == pu
                labels, core, st = ob.orc_labels_from_sim(rp, adj, r.sim, mu)
                assert np.array_equal(r.core, core), "core flags differ"
                diff = np.nonzero(r.labels != labels)[0]
                assert len(diff) == 0, f"labels differ at {diff[:10]}"
                for k in ("cores", "clusters", "borders", "hubs", "outliers", "similar_edges"):
                    assert getattr(r.stats, k) == st[k], k
                # the same pass without flags, with the pruned core check
                ctx.set_pruning(True)
                try:
                    p = sgc.scan(ctx, dg, eps, mu)
                finally:
                    ctx.set_pruning(False)
                assert p.stats.pruned == 1
                assert np.array_equal(p.core, r.core) and np.array_equal(p.labels, r.labels), "pruned pass differs"
                say(f"  pruned: {p.stats.deferred_edges} edges never decided, {p.stats.redo_edges} decided on demand")
                out[(eps, mu)] = st
    finally:
        dg.close()
    return out


def test_rmat24_grid(ctx):
    """R-MAT scale 24, edge factor 16, a/b/c/d .57/.19/.19/.05, seed 1: all
    twelve (eps, mu) passes of the benchmark grid."""
    g = sgc.HostGraph.rmat(24, 16, seed=1)
    say(f"R-MAT-24: n={g.n} m={g.m} max degree {g.max_degree}")
    assert g.max_degree > WRAP_ROW
    st = check_grid(ctx, g, [(eps, (2, 5, 10)) for eps in (0.2, 0.4, 0.6, 0.8)], seed=24)
    assert st[(0.2, 2)]["clusters"] > 1000  # the grid exercises clustering


def test_planted_partition_full(ctx):
    """Planted partition at its configured size: n = 2M, 20,000 communities
    with Zipf(1.5) sizes, p_in 0.05 capped at 200 expected intra-degree, 8
    inter-community edges per vertex, 64 hubs each joined to 5 % of the
    vertices; eps 0.5 mu 4 (no similar edge), and eps 0.08 / 0.05 mu 4."""
    g = sgc.HostGraph.planted(n=2_000_000, communities=20_000, zipf=1.5, p_in=0.05, max_intra_degree=200.0,
                              inter_per_vertex=8, hubs=64, hub_fraction=0.05, seed=1)
    say(f"planted: n={g.n} m={g.m} max degree {g.max_degree}")
    st = check_grid(ctx, g, [(0.5, (4,)), (0.08, (4,)), (0.05, (4,))], seed=3)
    assert st[(0.5, 4)]["similar_edges"] == 0
    assert st[(0.05, 4)]["clusters"] > 0 and st[(0.05, 4)]["borders"] > 0


def test_rmat24_sparse_without_window_index(ctx):
    """R-MAT-24 at edge factor 2 (n = 2^24, 16 windows of 2^20 ids, m = 67M):
    the window-offset index would take n * 17 * 4 = 1.14 GB, more than
    max(16 m, 1 GiB), so the resident pass runs without it and the
    workgroup tier finds each long owner row's window boundaries itself
    (hubs of degree up to 75k span every window).  Sampled flags, labels and
    core flags against the oracle, as above."""
    g = sgc.HostGraph.rmat(24, 2, seed=3)
    say(f"R-MAT-24 ef2: n={g.n} m={g.m} max degree {g.max_degree}")
    assert g.n * 17 * 4 > max(16 * g.m, 1 << 30) and g.max_degree > 8192
    check_grid(ctx, g, [(0.2, (2,)), (0.5, (3,))], seed=2)
