"""Autotuner of the similarity planner: any knob setting gives identical
results; the tuned setting is one of the candidates and no slower than the
defaults as measured (needs an MI355X)."""
from __future__ import annotations

import numpy as np
import pytest

import efficientstructuralgraphclusteringongpus_amd as sgc

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("knobs", [(128, 16384.0, 262144.0, 1, 512), (512, 262144.0, 4194304.0, 0, 0),
                                   (1024, 4096.0, 65536.0, 0, 65536), (1, 64.0, 64.0, 1, 2048)])
def test_any_knobs_give_identical_results(ctx, knobs):
    say("knobs", knobs)
    g = sgc.HostGraph.rmat(16, 16, seed=21)
    dg = sgc.DeviceGraph(ctx, g)
    try:
        ref = sgc.scan(ctx, dg, 0.3, 3, want_sim=True)
        dg.tuning = sgc.Tuning(*knobs)
        t = dg.tuning
        assert (t.small_max_degree, t.small_unit_work, t.large_unit_work, t.hash_tables, t.tiny_work) == knobs
        r = sgc.scan(ctx, dg, 0.3, 3, want_sim=True)
        assert np.array_equal(r.sim, ref.sim) and np.array_equal(r.labels, ref.labels)
    finally:
        dg.close()


@pytest.mark.parametrize("eps", [0.2, 0.5])
def test_hash_tables_off_over_several_windows(ctx, eps):
    """n > 2^20: the workgroup tier's cuckoo hashes (whole rows, window runs)
    and the small tier's against bitmap windows and linear probing only, and
    tiny owners merged by one thread against their own work units."""
    say("hash tables off", eps)
    g = sgc.HostGraph.rmat(21, 16, seed=12)
    dg = sgc.DeviceGraph(ctx, g)
    try:
        ref = sgc.scan(ctx, dg, eps, 3, want_sim=True)
        dg.tuning = sgc.Tuning(hash_tables=0)
        r = sgc.scan(ctx, dg, eps, 3, want_sim=True)
        assert np.array_equal(r.sim, ref.sim) and np.array_equal(r.labels, ref.labels)
        dg.tuning = sgc.Tuning(tiny_work=0)
        r = sgc.scan(ctx, dg, eps, 3, want_sim=True)
        assert np.array_equal(r.sim, ref.sim) and np.array_equal(r.labels, ref.labels)
    finally:
        dg.close()


def test_autotune_keeps_a_candidate_and_results(ctx):
    say("autotune")
    g = sgc.HostGraph.planted(n=200_000, communities=2_000, hubs=16, hub_fraction=0.05, seed=4)
    dg = sgc.DeviceGraph(ctx, g)
    try:
        ref = sgc.scan(ctx, dg, 0.5, 4, want_sim=True)
        best = dg.autotune(0.5)
        say("tuned", best)
        assert best.small_max_degree in (128, 256, 512, 1024)
        assert best.small_unit_work in (65536.0, 262144.0, 1048576.0)
        assert best.large_unit_work in (2097152.0, 8388608.0, 33554432.0)
        assert best.hash_tables in (0, 1)
        assert best.tiny_work in (0, 1024, 2048, 4096)
        assert best.trials >= 12 and best.ms_similarity <= best.ms_default
        t = dg.tuning
        assert (t.small_max_degree, t.large_unit_work) == (best.small_max_degree, best.large_unit_work)
        r = sgc.scan(ctx, dg, 0.5, 4, want_sim=True)
        assert np.array_equal(r.sim, ref.sim) and np.array_equal(r.labels, ref.labels)
    finally:
        dg.close()


def test_tuning_rejects_out_of_range_knobs(ctx):
    g = sgc.HostGraph.rmat(10, 8, seed=1)
    dg = sgc.DeviceGraph(ctx, g)
    try:
        for bad in ((0, 65536.0, 1048576.0), (2048, 65536.0, 1048576.0), (512, 1.0, 1048576.0),
                    (512, 65536.0, 1048576.0, 2), (512, 65536.0, 1048576.0, 1, 70000)):
            with pytest.raises(sgc.SGCError):
                dg.tuning = sgc.Tuning(*bad)
    finally:
        dg.close()
