"""The C++ drop-in int* GPUScan(graph_t*, int*, int*, double, int)
(GPUSCAN++/GPUScan.h:7) and its C twin sgc_GPUScan.

- CPU: libsgc.so exports the reference's mangled C++ symbol; an invalid
  graph_t makes either entry print the reference's "GPUassert: ..." line and
  exit non-zero (GPUSCAN++/GPUScan.cu:24-32), before any device call.
- GPU: a caller written like GPUSCAN++/main.cpp (tests/compat/gpuscan_caller)
  and the reference's own main.cpp + Graph.cpp, compiled unedited and linked
  against libsgc.so (oracle/_ref/GPUScan_refmain, built where the reference
  tree is present), print the oracle's core and cluster counts, and the
  caller's labels equal the oracle's.
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as ob

import efficientstructuralgraphclusteringongpus_amd as sgc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "compat", "_bin", "gpuscan_caller")
REFMAIN = os.path.join(ROOT, "oracle", "_ref", "GPUScan_refmain")
MANGLED = "_Z7GPUScanP7graph_tPiS1_di"  # GPUScan(graph_t*, int*, int*, double, int)


def say(*a):
    print("[case]", *a, flush=True)


def need_caller():
    if not os.path.exists(CALLER):
        pytest.fail(f"{CALLER} is not built (run __graft_entry__.build())")


def test_library_exports_reference_cxx_symbol():
    import ctypes as C

    lib = C.CDLL(sgc.library_path())
    assert hasattr(lib, MANGLED)
    assert hasattr(lib, "sgc_GPUScan")


@pytest.mark.parametrize("kind", ["n", "null-offsets", "offsets", "self-loop"])
def test_invalid_graph_t_prints_gpuassert_and_exits(tmp_path, kind):
    need_caller()
    g = sgc.HostGraph.rmat(8, 8, seed=3)
    folder = str(tmp_path / "g")
    g.save(folder)
    proc = subprocess.run([CALLER, folder, "0.5", "3", "--bad", kind], capture_output=True, text=True, timeout=120)
    assert proc.returncode != 0
    assert "GPUassert: " in proc.stderr, proc.stderr
    assert "Finish GPUScan" not in proc.stderr
    # the reference's lines before the call are unchanged
    assert f"node number and edge number: {g.n}, {g.m}" in proc.stderr


def test_c_entry_exits_on_invalid_graph():
    """sgc_GPUScan (C linkage) has the same contract; run in a child process."""
    code = (
        "import ctypes as C, sys\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "import efficientstructuralgraphclusteringongpus_amd as sgc\n"
        "lib = sgc.load_library()\n"
        "gt = sgc._GraphT(-5, 0, None, None, None)\n"
        "lib.sgc_GPUScan(C.byref(gt), None, None, 0.5, 3)\n"
        "print('returned')\n"
    )
    proc = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 1  # SGC_EINVAL
    assert "GPUassert: graph_t has n=-5" in proc.stderr
    assert "returned" not in proc.stdout


@pytest.mark.gpu
def test_cxx_caller_matches_oracle(tmp_path):
    need_caller()
    for scale, ef, seed, eps, mu in [(13, 16, 6, 0.5, 3), (12, 8, 2, 0.3, 5)]:
        say("gpuscan_caller", scale, ef, seed, eps, mu)
        g = sgc.HostGraph.rmat(scale, ef, seed=seed)
        folder = str(tmp_path / f"g{scale}")
        g.save(folder)
        out = str(tmp_path / f"labels{scale}.bin")
        proc = subprocess.run([CALLER, folder, str(eps), str(mu), "--labels", out], capture_output=True, text=True,
                              timeout=120)
        assert proc.returncode == 0, proc.stderr
        labels, core, _, st = ob.orc_scan(g.row_ptr, g.adj, eps, mu)
        assert np.array_equal(np.fromfile(out, dtype=np.int32), labels)
        assert f"core number: {st['cores']}" in proc.stderr
        assert f"Cluster Number: {st['clusters']}" in proc.stderr
        assert f"max degree: {g.max_degree}" in proc.stdout


@pytest.mark.gpu
def test_reference_main_links_against_libsgc(tmp_path):
    if not os.path.exists(REFMAIN):
        pytest.skip("oracle/_ref/GPUScan_refmain is built only where the reference tree is present")
    g = sgc.HostGraph.rmat(13, 16, seed=4)
    folder = str(tmp_path / "g")
    g.save(folder)
    for eps, mu in [(0.5, 3), (0.25, 4)]:
        say("reference main.cpp over libsgc", eps, mu)
        proc = subprocess.run([REFMAIN, folder, str(eps), str(mu)], capture_output=True, text=True, timeout=120)
        assert proc.returncode == 0, proc.stderr
        _, _, _, st = ob.orc_scan(g.row_ptr, g.adj, eps, mu)
        assert f"core number: {st['cores']}" in proc.stderr
        assert f"Cluster Number: {st['clusters']}" in proc.stderr
        assert "Work Complete!" in proc.stderr
