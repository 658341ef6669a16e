"""Per-vertex result files and the summary (host code, no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

import bruteforce as bf
import oracle_bind as ob

import efficientstructuralgraphclusteringongpus_amd as sgc


def oracle_result(n, edges, eps, mu):
    rp, adj = bf.csr(n, edges)
    labels, core, sim, st = ob.orc_scan(np.array(rp, dtype=np.uint64), np.array(adj, dtype=np.int32), eps, mu)
    stats = sgc.ScanStats(*([0.0] * 8), len(adj) // 2, st["prep_u"], st["similar_edges"], st["cores"],
                          st["clusters"], st["borders"], st["hubs"], st["outliers"], 0.0, 0.0, 0.0, 0.0)
    return sgc.ScanResult(labels, core, sim, stats), st


def test_write_and_read_back(tmp_path):
    # two triangles joined through a hub-like vertex, plus an isolated vertex
    edges = [(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5), (2, 6), (6, 3), (7, 8)]
    r, st = oracle_result(10, edges, 0.6, 2)
    out = str(tmp_path / "res")
    sgc.write_result(out, r, 0.6, 2, text=True)
    back = sgc.read_result(out)
    assert np.array_equal(back["labels"], r.labels) and np.array_equal(back["core"], r.core)
    role = back["role"]
    assert np.array_equal(role == 0, r.core.astype(bool))
    assert np.array_equal(role == 1, (~r.core.astype(bool)) & (r.labels >= 0))
    assert np.array_equal(role == 2, r.labels == -1) and np.array_equal(role == 3, r.labels == -2)
    s = back["summary"]
    assert s["core number"] == str(st["cores"]) and s["Cluster Number"] == str(st["clusters"])
    assert s["hubs"] == str(st["hubs"]) and s["outliers"] == str(st["outliers"]) and s["mu"] == "2"
    assert float(s["eps"]) == 0.6
    lines = open(tmp_path / "res" / "labels.txt").read().splitlines()
    assert lines[0].startswith("#") and len(lines) == 11
    v, lab, name = lines[1 + 9].split()
    assert int(v) == 9 and int(lab) == r.labels[9] and name == sgc.ROLE_NAMES[role[9]]


def test_summary_lines_match_the_reference_wording(tmp_path):
    r, st = oracle_result(6, [(0, 1), (1, 2), (0, 2), (2, 3), (3, 4), (4, 5), (3, 5)], 0.5, 2)
    sgc.write_result(str(tmp_path / "o"), r, 0.5, 2)
    text = open(tmp_path / "o" / "summary.txt").read().splitlines()
    assert text[0] == f"core number: {st['cores']}" and text[1] == f"Cluster Number: {st['clusters']}"
    assert not (tmp_path / "o" / "labels.txt").exists()


def test_rejects_labels_outside_the_encoding(tmp_path):
    r, _ = oracle_result(3, [(0, 1)], 0.5, 2)
    r.labels[1] = -7
    with pytest.raises(sgc.SGCError):
        sgc.write_result(str(tmp_path / "bad"), r, 0.5, 2)
