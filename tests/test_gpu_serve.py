"""Serving engine: every served query equals a one-shot SCAN pass of the same
graph, alone or mixed with other graphs' queries (needs an MI355X)."""
from __future__ import annotations

import random

import numpy as np
import pytest

import oracle_bind as ob

import efficientstructuralgraphclusteringongpus_amd as sgc

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def one_shot(ctx, dg, eps, mu):
    """A plain pass that decides every edge (flags requested), so that its
    similar-edge count is complete; a pruned pass counts decided edges only."""
    return sgc.scan(ctx, dg, eps, mu, want_sim=True)


def drain(srv, tickets):
    out = {}
    while len(out) < len(tickets):
        srv.progress(wait=True)
        for t, n in tickets.items():
            if t not in out:
                r = srv.result(t, n)
                if r is not None:
                    out[t] = r
    return out


@pytest.fixture(scope="module")
def graphs():
    return [sgc.HostGraph.rmat(14, 16, seed=11), sgc.HostGraph.rmat(16, 8, seed=12),
            sgc.HostGraph.planted(n=50_000, communities=500, hubs=4, hub_fraction=0.1, seed=13)]


def test_served_queries_match_one_shot_passes(ctx, graphs):
    rng = random.Random(3)
    with sgc.Server(ctx, 8 << 30, slots_per_graph=2) as srv:
        ids = [srv.add_graph(g)[0] for g in graphs]
        dgs = [sgc.DeviceGraph(ctx, g) for g in graphs]
        try:
            queries = {}
            for _ in range(48):
                i = rng.randrange(len(graphs))
                eps = rng.choice([0.1, 0.2, 0.3, 0.5, 0.7, 0.9])
                mu = rng.randint(2, 20)
                t = srv.submit(ids[i], eps, mu)
                queries[t] = (i, eps, mu)
            say("served", len(queries))
            res = drain(srv, {t: graphs[q[0]].n for t, q in queries.items()})
            for t, (i, eps, mu) in queries.items():
                labels, st = res[t]
                ref = one_shot(ctx, dgs[i], eps, mu)
                assert np.array_equal(labels, ref.labels), (i, eps, mu)
                assert (st.cores, st.clusters, st.borders, st.hubs, st.outliers, st.similar_edges) == (
                    ref.stats.cores, ref.stats.clusters, ref.stats.borders, ref.stats.hubs, ref.stats.outliers,
                    ref.stats.similar_edges), (i, eps, mu)
                assert st.ms_device > 0 and st.ms_latency > 0
        finally:
            for d in dgs:
                d.close()
        assert srv.pending == 0


def test_served_small_graph_against_oracle(ctx):
    g = sgc.HostGraph.rmat(11, 8, seed=2)
    rp, adj = np.array(g.row_ptr), np.array(g.adj)
    with sgc.Server(ctx, 1 << 30, slots_per_graph=3) as srv:
        gid, ms = srv.add_graph(g)
        assert ms >= 0
        for eps, mu in ((0.2, 2), (0.35, 3), (0.6, 2), (1.0, 1), (0.0, 0)):
            say("served vs oracle", eps, mu)
            t = srv.submit(gid, eps, mu)
            labels, _ = drain(srv, {t: g.n})[t]
            ref, _, _, _ = ob.orc_scan(rp, adj, eps, mu)
            assert np.array_equal(labels, ref)


def test_mixed_load_gives_the_same_answers_as_alone(ctx, graphs):
    a, b = graphs[0], graphs[1]
    with sgc.Server(ctx, 8 << 30, slots_per_graph=4) as srv:
        ga, _ = srv.add_graph(a)
        t = srv.submit(ga, 0.3, 3)
        alone = drain(srv, {t: a.n})[t][0]
        gb, _ = srv.add_graph(b)
        tickets = {}
        for k in range(12):
            tickets[srv.submit(gb, 0.2 + 0.05 * (k % 5), 2 + k % 4)] = b.n
            tickets[srv.submit(ga, 0.3, 3)] = a.n
        res = drain(srv, tickets)
        for tk, n in tickets.items():
            if n == a.n:
                assert np.array_equal(res[tk][0], alone)


def test_admission_and_removal(ctx, graphs):
    g = graphs[1]
    need = sgc.Server.graph_bytes(g, 2)
    with sgc.Server(ctx, need + need // 2, slots_per_graph=2) as srv:
        gid, _ = srv.add_graph(g)
        assert srv.bytes_in_use == need
        with pytest.raises(sgc.SGCError) as e:
            srv.add_graph(g)
        assert sgc.SGCError.CODES[e.value.code] == "ENOMEM"
        t = srv.submit(gid, 0.4, 2)
        srv.remove_graph(gid)  # waits for the running query, keeps its result
        assert srv.bytes_in_use == 0
        drain(srv, {t: g.n})
        gid2, _ = srv.add_graph(g)
        t2 = srv.submit(gid2, 0.4, 2)
        assert drain(srv, {t2: g.n})[t2][0].shape == (g.n,)
        with pytest.raises(sgc.SGCError):
            srv.submit(gid, 0.4, 2)  # removed graph


@pytest.mark.parametrize("batched", [False, True])
def test_removal_fails_queued_queries(ctx, graphs, batched):
    """A graph removed with queries still queued: the running ones finish
    with their results, the queued ones end with an error result (EINVAL)
    instead of running on freed memory; nothing is left open."""
    g = graphs[1]
    dg = sgc.DeviceGraph(ctx, g)
    try:
        ref = {eps: one_shot(ctx, dg, eps, 3).labels for eps in (0.3, 0.5)}
    finally:
        dg.close()
    with sgc.Server(ctx, 4 << 30, slots_per_graph=1) as srv:
        gid, _ = srv.add_graph(g)
        if batched:
            srv.set_batching(1 << 40, 1)
        # one slot and ~1 ms per query on the GPU: most of these are still
        # queued when the graph goes
        tickets = {srv.submit(gid, 0.3 if i % 2 else 0.5, 3): 0.3 if i % 2 else 0.5 for i in range(48)}
        srv.remove_graph(gid)
        ok = failed = 0
        for _ in range(3):
            srv.progress(wait=False)
        # every ticket is listed as completed, failed ones included
        assert set(srv.completed()) == set(tickets)
        for t, eps in tickets.items():
            try:
                r = srv.result(t, g.n)
            except sgc.SGCError as e:
                assert sgc.SGCError.CODES[e.code] == "EINVAL" and "removed graph" in e.message
                failed += 1
                continue
            assert r is not None, "a query of a removed graph is still open"
            assert np.array_equal(r[0], ref[eps])
            ok += 1
        say("removed with queries queued", "batched" if batched else "slots", ok, "finished", failed, "failed")
        assert ok >= 1 and failed >= 1 and srv.pending == 0


def test_edgeless_graph_served(ctx):
    g = sgc.HostGraph.from_csr([0, 0, 0, 0], [])
    with sgc.Server(ctx, 1 << 26, slots_per_graph=1) as srv:
        gid, _ = srv.add_graph(g)
        t = srv.submit(gid, 0.5, 2)
        labels, st = drain(srv, {t: 3})[t]
        assert list(labels) == [-2, -2, -2] and st.outliers == 3


def test_graph_replays_match_eager_launches(ctx, graphs):
    """One slot, so every query replays the same captured hipGraph; eager work
    and synchronisation run between the replays and every replay is checked
    against the eager launch mode and a one-shot pass."""
    g = graphs[1]
    dg = sgc.DeviceGraph(ctx, g)
    try:
        params = [(0.2, 2), (0.5, 5), (0.3, 3), (0.9, 2), (0.3, 3), (0.1, 10)]
        results = {}
        for mode in (True, False):
            with sgc.Server(ctx, 4 << 30, slots_per_graph=1) as srv:
                srv.set_graph_launch(mode)
                gid, _ = srv.add_graph(g)
                for eps, mu in params:
                    say("replay", mode, eps, mu)
                    t = srv.submit(gid, eps, mu)
                    labels, st = drain(srv, {t: g.n})[t]
                    ref = one_shot(ctx, dg, eps, mu)  # eager work between replays
                    ctx.sync()
                    assert np.array_equal(labels, ref.labels), (mode, eps, mu)
                    assert st.cores == ref.stats.cores and st.similar_edges == ref.stats.similar_edges
                    results[(mode, eps, mu)] = labels
        for eps, mu in params:
            assert np.array_equal(results[(True, eps, mu)], results[(False, eps, mu)])
    finally:
        dg.close()


def test_completed_lists_each_retired_ticket_once(ctx, graphs):
    """sgc_server_completed lists every retired query exactly once, and the
    results collected through it are the ones polling each ticket gives."""
    say("completed tickets")
    with sgc.Server(ctx, 1 << 32, slots_per_graph=2) as srv:
        gids = [srv.add_graph(g)[0] for g in graphs]
        ref = {}
        tickets = {}
        for k in range(24):
            gi = k % len(graphs)
            t = srv.submit(gids[gi], 0.2 + 0.1 * (k % 6), 2 + k % 5)
            tickets[t] = (gi, 0.2 + 0.1 * (k % 6), 2 + k % 5)
        seen = []
        got = {}
        while len(got) < len(tickets):
            srv.progress(wait=True)
            for t in srv.completed():
                seen.append(t)
                gi = tickets[t][0]
                got[t] = srv.result(t, graphs[gi].n)
        assert sorted(seen) == sorted(tickets) and len(seen) == len(set(seen))
        assert srv.completed() == []
        for t, (gi, eps, mu) in tickets.items():
            dg = sgc.DeviceGraph(ctx, graphs[gi])
            try:
                one = sgc.scan(ctx, dg, eps, mu)
            finally:
                dg.close()
            assert np.array_equal(got[t][0], one.labels)
