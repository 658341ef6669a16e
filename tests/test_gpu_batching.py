"""Batched serving (one launch per phase over a table of queries) against
one-shot plain passes, through the C ABI (needs an MI355X).

Every batched query must give the labels and counts of a one-shot pass
(GPUSCAN++/GPUScan.cu:800-1009 semantics, every edge decided), whether it ran
alone, in a batch with queries on other graphs, or beside queries replayed
one per slot.
"""
from __future__ import annotations

import random

import numpy as np
import pytest

import efficientstructuralgraphclusteringongpus_amd as sgc

pytestmark = pytest.mark.gpu

FIELDS = ("cores", "clusters", "borders", "hubs", "outliers", "similar_edges")


def say(*a):
    print("[case]", *a, flush=True)


def drain(srv, tickets):
    """tickets: {ticket: n}; returns {ticket: (labels, stats)}."""
    out = {}
    while len(out) < len(tickets):
        srv.progress(wait=True)
        for t in srv.completed():
            if t in tickets and t not in out:
                out[t] = srv.result(t, tickets[t])
    return out


@pytest.fixture(scope="module")
def small_graphs():
    return [sgc.HostGraph.rmat(11, 8, seed=21), sgc.HostGraph.rmat(12, 16, seed=22),
            sgc.HostGraph.planted(n=20_000, communities=200, hubs=2, hub_fraction=0.1, seed=23),
            sgc.HostGraph.from_csr([0, 0, 0], []),  # edgeless
            sgc.HostGraph.from_csr([0, 1, 2], [1, 0])]  # one edge


def one_shot_all(ctx, graphs, queries):
    """{(graph index, eps, mu): ScanResult} of one-shot passes deciding every edge."""
    ref = {}
    dgs = {}
    try:
        for gi, eps, mu in set(queries):
            if gi not in dgs:
                dgs[gi] = sgc.DeviceGraph(ctx, graphs[gi])
            ref[(gi, eps, mu)] = sgc.scan(ctx, dgs[gi], eps, mu, want_sim=True)
    finally:
        for d in dgs.values():
            d.close()
    return ref


def check(results, tickets, ref):
    for t, key in tickets.items():
        labels, st = results[t]
        r = ref[key]
        assert np.array_equal(labels, r.labels), key
        assert tuple(getattr(st, f) for f in FIELDS) == tuple(getattr(r.stats, f) for f in FIELDS), key


def test_batched_queries_match_one_shot(ctx, small_graphs):
    rng = random.Random(7)
    queries = [(rng.randrange(len(small_graphs)), rng.choice([0.1, 0.3, 0.5, 0.7]), rng.choice([-1, 0, 2, 3, 6]))
               for _ in range(96)]
    ref = one_shot_all(ctx, small_graphs, queries)
    with sgc.Server(ctx, 4 << 30, slots_per_graph=16) as srv:
        ids = [srv.add_graph(g)[0] for g in small_graphs]
        srv.set_batching(1 << 40, 64)  # every graph batched
        assert srv.batching() == (1 << 40, 64)
        tickets = {}
        for gi, eps, mu in queries:
            tickets[srv.submit(ids[gi], eps, mu)] = (gi, eps, mu)
        say("batched", len(tickets))
        res = drain(srv, {t: small_graphs[k[0]].n for t, k in tickets.items()})
        check(res, tickets, ref)
        assert srv.pending == 0


def test_batch_cap_and_many_batches(ctx, small_graphs):
    """More queued queries than one batch holds (cap 5), all on one graph with
    64 slots: several batches in a row, two lanes at once."""
    g = small_graphs[1]
    with sgc.Server(ctx, 4 << 30, slots_per_graph=64) as srv:
        gid, _ = srv.add_graph(g)
        srv.set_batching(1 << 40, 5)
        params = [(0.2 + 0.05 * (i % 10), 2 + i % 4) for i in range(150)]
        tickets = {srv.submit(gid, eps, mu): (0, eps, mu) for eps, mu in params}
        say("many batches", len(tickets))
        res = drain(srv, {t: g.n for t in tickets})
    check(res, tickets, one_shot_all(ctx, [g], list(tickets.values())))


def test_batches_beside_slot_replays(ctx, small_graphs):
    """Cut-off between graph sizes: the small graphs' queries run in batches,
    the larger graph's in slots, at the same time; a query gives the same
    result alone and in the mix."""
    graphs = small_graphs[:3] + [sgc.HostGraph.rmat(15, 16, seed=24)]
    cut = max(g.m for g in graphs[:2])  # the two small R-MAT graphs batched
    assert sum(g.m > cut for g in graphs) == 2  # the planted graph and R-MAT-15 replayed
    rng = random.Random(9)
    queries = [(rng.randrange(len(graphs)), rng.choice([0.2, 0.4, 0.6]), rng.choice([2, 4])) for _ in range(64)]
    ref = one_shot_all(ctx, graphs, queries)
    with sgc.Server(ctx, 8 << 30, slots_per_graph=8) as srv:
        ids = [srv.add_graph(g)[0] for g in graphs]
        srv.set_batching(cut, 64)
        alone = {}
        for key in [(0, 0.2, 2), (3, 0.4, 4)]:
            t = srv.submit(ids[key[0]], key[1], key[2])
            alone[key] = drain(srv, {t: graphs[key[0]].n})[t][0]
        tickets = {srv.submit(ids[gi], eps, mu): (gi, eps, mu) for gi, eps, mu in queries}
        say("mixed", len(tickets))
        res = drain(srv, {t: graphs[k[0]].n for t, k in tickets.items()})
    check(res, tickets, ref)
    ref_alone = one_shot_all(ctx, graphs, list(alone))
    for key, labels in alone.items():
        assert np.array_equal(labels, ref_alone[key].labels), key


def test_autotuned_cutoff(ctx, small_graphs):
    graphs = small_graphs[:3] + [sgc.HostGraph.rmat(16, 16, seed=25)]
    with sgc.Server(ctx, 8 << 30, slots_per_graph=4) as srv:
        ids = [srv.add_graph(g)[0] for g in graphs]
        cut = srv.autotune_batching(0.5, 5)
        say("autotuned cut-off", cut, [g.m for g in graphs])
        assert cut == 0 or cut in [g.m for g in graphs]
        assert srv.batching()[0] == cut
        rng = random.Random(11)
        queries = [(rng.randrange(len(graphs)), rng.choice([0.3, 0.5]), rng.choice([2, 5])) for _ in range(32)]
        tickets = {srv.submit(ids[gi], eps, mu): (gi, eps, mu) for gi, eps, mu in queries}
        res = drain(srv, {t: graphs[k[0]].n for t, k in tickets.items()})
    check(res, tickets, one_shot_all(ctx, graphs, queries))
