"""Input surface: b_degree.bin/b_adj.bin reader/writer and edge orientation.

Checked against byte layouts written by hand, against the reference's own
Graph::Graph and getEidAndEdgeList (compiled from /root/reference into
oracle/_ref/ where the reference tree is present), and against the oracle.
"""
from __future__ import annotations

import os
import random
import struct
import subprocess
import sys

import numpy as np
import pytest

import bruteforce as bf
import oracle_bind as ob
import efficientstructuralgraphclusteringongpus_amd as sgc

needs_ref = pytest.mark.skipif(not ob.reference_available(), reason="reference host code not built here")


def write_raw(folder, n, degs, adj, int_size=4, m_field=None):
    os.makedirs(folder, exist_ok=True)
    m = sum(degs) if m_field is None else m_field
    with open(os.path.join(folder, "b_degree.bin"), "wb") as f:
        f.write(struct.pack("<iII", int_size, n, m & 0xFFFFFFFF))
        f.write(struct.pack("<%di" % len(degs), *degs))
    with open(os.path.join(folder, "b_adj.bin"), "wb") as f:
        f.write(struct.pack("<%di" % len(adj), *adj))


def test_load_hand_written_bytes(tmp_path):
    # triangle 0-1-2 plus isolated vertex 3
    write_raw(str(tmp_path), 4, [2, 2, 2, 0], [1, 2, 0, 2, 0, 1])
    g = sgc.HostGraph.load(str(tmp_path), check_symmetry=True)
    assert g.n == 4 and g.m == 6
    assert list(g.row_ptr) == [0, 2, 4, 6, 6]
    assert list(g.adj) == [1, 2, 0, 2, 0, 1]
    assert g.max_degree == 2


def test_save_writes_reference_layout(tmp_path):
    rp, adj = bf.csr(5, [(0, 1), (1, 2), (3, 4), (0, 4)])
    g = sgc.HostGraph.from_csr(rp, adj)
    g.save(str(tmp_path))
    raw = open(os.path.join(tmp_path, "b_degree.bin"), "rb").read()
    assert struct.unpack("<iII", raw[:12]) == (4, 5, len(adj))
    assert list(struct.unpack("<5i", raw[12:])) == [2, 2, 1, 1, 2]
    assert open(os.path.join(tmp_path, "b_adj.bin"), "rb").read() == struct.pack("<%di" % len(adj), *adj)
    g2 = sgc.HostGraph.load(str(tmp_path))
    assert g2.digest() == g.digest()


@pytest.mark.parametrize("case,code,needle", [
    ("missing", 2, "cannot open"),
    ("truncated_header", 2, "truncated header"),
    ("int_size", 3, "int_size"),
    ("m_mismatch", 3, "does not match"),
    ("short_adj", 2, "bytes, expected"),
    ("long_adj", 3, "bytes, expected"),
    ("self_loop", 3, "Self loop"),
    ("unsorted", 3, "not sorted"),
    ("duplicate", 3, "not sorted"),
    ("out_of_range", 3, "out of range"),
    ("negative", 3, "out of range"),
])
def test_load_rejects(tmp_path, case, code, needle):
    d = str(tmp_path / case)
    if case == "missing":
        os.makedirs(d)
    elif case == "truncated_header":
        os.makedirs(d)
        open(os.path.join(d, "b_degree.bin"), "wb").write(b"\x04\x00\x00\x00\x03")
    elif case == "int_size":
        write_raw(d, 2, [1, 1], [1, 0], int_size=8)
    elif case == "m_mismatch":
        write_raw(d, 2, [1, 1], [1, 0], m_field=4)
    elif case == "short_adj":
        write_raw(d, 2, [1, 1], [1])
        with open(os.path.join(d, "b_degree.bin"), "r+b") as f:
            f.seek(8)
            f.write(struct.pack("<I", 2))
    elif case == "long_adj":
        write_raw(d, 2, [1, 1], [1, 0, 1])  # one entry beyond the degree sum
    elif case == "self_loop":
        write_raw(d, 2, [2, 1], [0, 1, 0])
    elif case == "unsorted":
        write_raw(d, 3, [2, 1, 1], [2, 1, 0, 0])
    elif case == "duplicate":
        write_raw(d, 2, [2, 2], [1, 1, 0, 0])
    elif case == "out_of_range":
        write_raw(d, 2, [1, 1], [5, 0])
    elif case == "negative":
        write_raw(d, 2, [1, 1], [-1, 0])
    with pytest.raises(sgc.SGCError) as ei:
        sgc.HostGraph.load(d)
    assert ei.value.code == code
    assert needle in ei.value.message


def test_asymmetric_rejected_when_checked(tmp_path):
    write_raw(str(tmp_path), 3, [1, 1, 0], [1, 2])
    sgc.HostGraph.load(str(tmp_path))  # the reference does not check symmetry either
    with pytest.raises(sgc.SGCError) as ei:
        sgc.HostGraph.load(str(tmp_path), check_symmetry=True)
    assert "symmetric" in ei.value.message


def test_edge_list_converter(tmp_path):
    txt = tmp_path / "g.txt"
    txt.write_text("# comment\n0 1\n1 0\n1 2\n2 2\n% other comment\n3 1\n0 1\n")
    out = str(tmp_path / "bin")
    os.makedirs(out)
    sgc.HostGraph.convert_edge_list(str(txt), out)
    g = sgc.HostGraph.load(out, check_symmetry=True)
    assert g.n == 4
    assert list(g.row_ptr) == [0, 1, 4, 5, 6]
    assert list(g.adj) == [1, 0, 2, 3, 1, 1]


def test_orient_matches_oracle_on_random_graphs():
    rng = random.Random(5)
    for trial in range(20):
        n = rng.randint(2, 200)
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(1, 6 * n))]
        rp, adj = bf.csr(n, edges)
        if not adj:
            continue
        g = sgc.HostGraph.from_csr(rp, adj)
        got = g.orient()
        exp = ob.orc_orient(rp, adj)
        for a, b in zip(got, exp):
            assert np.array_equal(a, b)


@needs_ref
def test_loader_matches_reference_graph_reader(tmp_path):
    g = sgc.HostGraph.rmat(12, 8, seed=3)
    g.save(str(tmp_path))
    rp, adj = ob.ref_load(str(tmp_path))
    assert np.array_equal(rp.astype(np.uint64), g.row_ptr)
    assert np.array_equal(adj, g.adj)
    g2 = sgc.HostGraph.load(str(tmp_path))
    assert np.array_equal(g2.row_ptr, g.row_ptr) and np.array_equal(g2.adj, g.adj)


@needs_ref
@pytest.mark.parametrize("scale,ef,seed", [(8, 4, 1), (11, 16, 2), (13, 8, 9)])
def test_orientation_matches_reference(scale, ef, seed):
    g = sgc.HostGraph.rmat(scale, ef, seed=seed)
    rp, adj = g.row_ptr.astype(np.uint32), np.array(g.adj)
    ref_eu, ref_ev, ref_eid = ob.ref_orient(rp, adj)
    eu, ev, eid = g.orient()
    assert np.array_equal(eu, ref_eu)
    assert np.array_equal(ev, ref_ev)
    assert np.array_equal(eid, ref_eid)
    o_eu, o_ev, o_eid = ob.orc_orient(g.row_ptr, adj)
    assert np.array_equal(o_eu, ref_eu) and np.array_equal(o_ev, ref_ev) and np.array_equal(o_eid, ref_eid)


@needs_ref
def test_reference_rejects_what_loader_rejects(tmp_path):
    """The reference exits(1) with "Self loop" / "not sorted" (Graph.cpp:276-283);
    the library returns SGC_EFORMAT with the same words."""
    cases = {"self_loop": ([2, 1], [0, 1, 0], "Self loop"), "unsorted": ([2, 1, 1], [2, 1, 0, 0], "not sorted")}
    for name, (degs, adj, needle) in cases.items():
        d = str(tmp_path / name)
        write_raw(d, len(degs), degs, adj)
        code = ("import sys; sys.path.insert(0, %r); import oracle_bind as ob; ob.ref_load(%r)"
                % (os.path.dirname(__file__), d))
        proc = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
        assert proc.returncode == 1 and needle in proc.stderr
        with pytest.raises(sgc.SGCError) as ei:
            sgc.HostGraph.load(d)
        assert needle in ei.value.message


@pytest.mark.parametrize("degs,m_field,code,needle", [
    # degree sum exactly 2^32: the header holds 0 (the 64-bit extension), so
    # the header is accepted and the loader asks for 2^34 bytes of b_adj.bin
    ([2**31 - 1, 2**31 - 1, 2], 0, 2, "expected 17179869184"),
    # degree sum 2^32 + 7: header 7
    ([2**31 - 1, 2**31 - 1, 9], 7, 2, "expected 17179869212"),
    # a header that is not the degree sum mod 2^32 is refused before any read
    ([2**31 - 1, 2**31 - 1, 2], 5, 3, "does not match the degree sum 4294967296"),
    # the reference's own reading (a 32-bit m equal to the sum) below 2^32
    ([2**31 - 1, 3], 2**31 + 2, 2, "expected 8589934600"),
])
def test_header_64bit_extension(tmp_path, degs, m_field, code, needle):
    """b_degree.bin's uint32 m field holds sum(deg) mod 2^32 when the sum
    reaches 2^32 (the reference writes m as 4 bytes,
    csr_convert/EdgeListToBinsConverter.cpp:103-107, and cannot represent
    it).  The header arithmetic is checked before b_adj.bin is read, so
    degree arrays summing past 2^32 with a short adjacency file pin it
    without a 16 GB file."""
    d = str(tmp_path / "g")
    write_raw(d, len(degs), degs, [1, 0], m_field=m_field)
    with pytest.raises(sgc.SGCError) as ei:
        sgc.HostGraph.load(d)
    assert ei.value.code == code, ei.value.message
    assert needle in ei.value.message, ei.value.message
