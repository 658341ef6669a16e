"""sgc_scan_device_noalloc (the pass without any device allocation) against
sgc_scan_device, through the C ABI: labels, core flags, similarity flags and
statistics bit for bit, over two calls in a row on the shared static
scratch, the second on a smaller graph; a graph whose pass needs more than
the scratch is refused with SGC_ERANGE.  Device output buffers come from
hipMalloc in the test, outside the calls under test."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob

import efficientstructuralgraphclusteringongpus_amd as sgc

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        sgc.load_library()  # libsgc.so brings the HIP runtime in
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return _hip


class DevOut:
    """Device buffers for one pass's labels, core flags and similarity flags."""

    def __init__(self, n, m):
        self.n, self.m = n, m
        self.ptrs = []
        for nbytes in (4 * max(n, 1), max(n, 1), max(m, 1)):
            p = C.c_void_p()
            assert hip().hipMalloc(C.byref(p), nbytes) == 0
            assert hip().hipMemset(p, 0xAB, nbytes) == 0  # stale bytes a pass must overwrite
            self.ptrs.append(p.value)

    def fetch(self):
        labels = np.empty(max(self.n, 1), np.int32)
        core = np.empty(max(self.n, 1), np.uint8)
        sim = np.empty(max(self.m, 1), np.uint8)
        for host, dev in zip((labels, core, sim), self.ptrs):
            assert hip().hipMemcpy(host.ctypes.data, dev, host.nbytes, 2) == 0  # device to host
        return labels[: self.n], core[: self.n], sim[: self.m]

    def free(self):
        for p in self.ptrs:
            hip().hipFree(p)


STAT_FIELDS = ("edges", "u_edges", "similar_edges", "cores", "clusters", "borders", "hubs", "outliers")


def test_noalloc_matches_scan_device(ctx):
    cases = [(sgc.HostGraph.rmat(14, 16, seed=5), 0.4, 3), (sgc.HostGraph.rmat(11, 8, seed=6), 0.3, 2)]
    dgs = [sgc.DeviceGraph(ctx, g) for g, _, _ in cases]
    try:
        for dg in dgs:
            assert sgc.noalloc_bytes(dg) <= sgc.noalloc_max_bytes()
        # two calls in a row on the shared scratch, the second on the smaller graph
        got = []
        for (g, eps, mu), dg in zip(cases, dgs):
            print(f"[case] noalloc n={g.n} m={g.m} eps={eps} mu={mu}", flush=True)
            out = DevOut(g.n, g.m)
            st = sgc.scan_device_noalloc(ctx, dg, eps, mu, *out.ptrs)
            got.append((out.fetch(), st))
            out.free()
        for ((g, eps, mu), dg), ((labels, core, sim), st) in zip(zip(cases, dgs), got):
            print(f"[case] scan_device n={g.n} m={g.m} eps={eps} mu={mu}", flush=True)
            out = DevOut(g.n, g.m)
            ref_st = sgc.scan_device(ctx, dg, eps, mu, *out.ptrs)
            rl, rc, rs = out.fetch()
            out.free()
            assert np.array_equal(labels, rl) and np.array_equal(core, rc) and np.array_equal(sim, rs)
            for k in STAT_FIELDS:
                assert getattr(st, k) == getattr(ref_st, k), k
            ol, oc, osim, _ = ob.orc_scan(g.row_ptr, g.adj, eps, mu)
            assert np.array_equal(labels, ol) and np.array_equal(core, oc) and np.array_equal(sim, osim)
    finally:
        for dg in dgs:
            dg.close()


def test_noalloc_refuses_graph_beyond_scratch(ctx):
    g = sgc.HostGraph.rmat(20, 16, seed=1)
    dg = sgc.DeviceGraph(ctx, g)
    try:
        assert sgc.noalloc_bytes(dg) > sgc.noalloc_max_bytes()
        print(f"[case] noalloc refusal n={g.n} m={g.m}", flush=True)
        with pytest.raises(sgc.SGCError) as ei:
            sgc.scan_device_noalloc(ctx, dg, 0.5, 3)
        assert ei.value.code == 6 and "static scratch" in ei.value.message  # SGC_ERANGE
        # the refused call left the context usable
        sgc.scan_device(ctx, dg, 0.5, 3)
    finally:
        dg.close()
