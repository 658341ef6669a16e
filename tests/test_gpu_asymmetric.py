"""Inputs without a reverse entry for every entry (the reference's format
allows them, and neither it nor the loader's default checks for them): a
pass whose planner would need more work units than a symmetric graph can
stops before the similarity kernels and reports SGC_EFORMAT (needs an
MI355X)."""
from __future__ import annotations

import numpy as np
import pytest

import efficientstructuralgraphclusteringongpus_amd as sgc

pytestmark = pytest.mark.gpu
EFORMAT = 3


def say(*a):
    print("[case]", *a, flush=True)


def downward_graph(n=1000, k=8):
    """Vertex u lists u-k .. u-1 only: every entry points to its lower
    (degree, id) endpoint, so at eps 0.5 every entry is owned and undecided,
    twice the entries a symmetric graph of the same size can own."""
    rp, adj = [0], []
    for u in range(n):
        adj.extend(range(max(0, u - k), u))
        rp.append(len(adj))
    return sgc.HostGraph.from_csr(np.array(rp), np.array(adj, dtype=np.int32), check_symmetry=False)


def test_resident_pass_reports_asymmetric_input(ctx):
    say("asymmetric resident")
    g = downward_graph()
    dg = sgc.DeviceGraph(ctx, g)
    try:
        with pytest.raises(sgc.SGCError) as ei:
            sgc.scan(ctx, dg, 0.5, 2)
        assert ei.value.code == EFORMAT and "not symmetric" in ei.value.message
        # the context stays usable: a symmetric graph still runs
        ok = sgc.HostGraph.rmat(12, 8, seed=2)
        dg2 = sgc.DeviceGraph(ctx, ok)
        r = sgc.scan(ctx, dg2, 0.5, 2)
        assert r.labels.shape == (ok.n,)
        dg2.close()
    finally:
        dg.close()


def test_out_of_core_pass_on_asymmetric_input(ctx):
    """A chip's unit array holds as many units as the chip's two blocks have
    entries over two, so a chip overflows only when its pair's cross entries
    outnumber that; here they do not, and the pass completes.  Either way it
    ends with results or SGC_EFORMAT, never with a fault."""
    say("asymmetric out of core")
    g = downward_graph()
    budget = sgc.ooc_min_budget(g, 0.5) + 37 * g.m
    try:
        r, info = sgc.scan_ooc(ctx, g, 0.5, 2, budget)
    except sgc.SGCError as e:
        assert e.code == EFORMAT
    else:
        assert r.labels.shape == (g.n,)


def test_server_rejects_asymmetric_graph(ctx):
    say("asymmetric serving")
    g = downward_graph()
    with sgc.Server(ctx, sgc.Server.graph_bytes(g, 2) + (64 << 20), slots_per_graph=2) as srv:
        with pytest.raises(sgc.SGCError) as ei:
            srv.add_graph(g)
        assert ei.value.code == EFORMAT
