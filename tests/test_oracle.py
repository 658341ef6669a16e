"""Pins the oracle (oracle/scan_oracle.c) before anything is checked against it.

* prep decision and similarity predicate: exact rational arithmetic wherever
  the double evaluation cannot round across the boundary, the reference's
  literal left-to-right expression at exact ties, and the int32 wrap of
  (cn+2)^2 at its boundary;
* common-neighbour counts: python set intersection;
* the whole pass: hand-computed known answers (tests/golden/kat_graphs.json),
  every graph on up to 5 vertices and random graphs against the declarative
  brute force (tests/bruteforce.py);
* orientation: the reference's own getEidAndEdgeList, compiled from its
  sources (tests/test_reference_host.py).
"""
from __future__ import annotations

import itertools
import json
import os
import random
from fractions import Fraction

import numpy as np
import pytest

import bruteforce as bf
import oracle_bind as ob

import efficientstructuralgraphclusteringongpus_amd as sgc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
EPS_GRID = [round(0.05 * i, 2) for i in range(1, 21)] + [0.0, 0.33, 1.25]


def exact_prep(da, db, eps):
    e = Fraction(eps)
    A, B = da + 1, db + 1
    if A < B * e * e:
        return 0
    if e * e * A * B <= 4:
        return 1
    return 2


def test_prep_matches_exact_rationals_off_ties():
    orc = ob.oracle()
    checked = ties = 0
    for eps in EPS_GRID:
        e2 = Fraction(eps) ** 2
        for da in range(0, 40):
            for db in range(da, 80):
                A, B = da + 1, db + 1
                near = abs(Fraction(A) - B * e2) <= Fraction(A, 2 ** 49) or abs(e2 * A * B - 4) <= Fraction(1, 2 ** 45)
                got = orc.orc_prep(da, db, eps)
                if near:
                    ties += 1
                    literal = 0 if float(A) < float(B) * eps * eps else (1 if eps * eps * float(A) * float(B) <= 4.0 else 2)
                    assert got == literal, (da, db, eps)
                else:
                    assert got == exact_prep(da, db, eps), (da, db, eps)
                checked += 1
    assert checked > 30000 and ties > 0


def exact_similar(cn, s, l, eps):
    lhs = bf.wrap_i32((cn + 2) * (cn + 2))
    return not (Fraction(lhs) < Fraction(eps) ** 2 * (s + 1) * (l + 1))


def test_predicate_matches_exact_rationals_off_ties():
    orc = ob.oracle()
    ties = 0
    for eps in EPS_GRID:
        e2 = Fraction(eps) ** 2
        for s in range(0, 48):
            for l in range(s, 64):
                rhs = e2 * (s + 1) * (l + 1)
                for cn in range(0, s + 1):
                    lhs = (cn + 2) ** 2
                    got = orc.orc_predicate(cn, s, l, eps)
                    if abs(lhs - rhs) <= rhs / 2 ** 48:
                        ties += 1
                        assert got == int(not (float(lhs) < eps * eps * float(s + 1) * float(l + 1)))
                    else:
                        assert got == int(exact_similar(cn, s, l, eps)), (cn, s, l, eps)
    assert ties > 0


@pytest.mark.parametrize("eps", [0.1, 0.2, 0.45, 0.9])
def test_predicate_operand_order_is_the_references(eps):
    """Where (cn+2)^2 sits on eps^2*A*B to within rounding, the decision depends
    on the evaluation order; the oracle must follow ((eps*eps)*A)*B
    (GPUSCAN++/GPUScan.cu:220-221).  The grid contains ties that the orders
    A*B*(eps*eps) and eps*(eps*(A*B)) decide the other way."""
    orc = ob.oracle()
    differs = 0
    for A in range(2, 300):
        for B in range(A, min(30 * A, 1500)):
            lit = eps * eps * float(A) * float(B)
            t = round(lit ** 0.5)
            if t < 2:
                continue
            lhs = float(t * t)
            literal = not (lhs < lit)
            alt1 = not (lhs < float(A) * float(B) * (eps * eps))
            alt2 = not (lhs < eps * (eps * (float(A) * float(B))))
            differs += (literal != alt1) or (literal != alt2)
            assert orc.orc_predicate(t - 2, A - 1, B - 1, eps) == int(literal), (A, B, t)
    assert differs > 0, "grid does not separate the evaluation orders"


@pytest.mark.parametrize("t", [46339, 46340, 46341, 46342, 65535, 65536, 65537, 92682, 92683])
def test_predicate_int32_wrap(t):
    orc = ob.oracle()
    cn = t - 2
    s = l = cn + 10
    for eps in (0.2, 0.5, 0.9, 1.0):
        lhs = bf.wrap_i32(t * t)
        expect = not (Fraction(lhs) < Fraction(eps) ** 2 * (s + 1) * (l + 1))
        assert orc.orc_predicate(cn, s, l, eps) == int(expect)
    if t >= 46341 and t < 65536:
        # wrapped negative: never similar however close the rows are
        assert orc.orc_predicate(cn, cn + 1, cn + 1, 0.01) == 0


def test_common_neighbours_against_sets():
    rng = random.Random(7)
    for trial in range(30):
        n = rng.randint(2, 60)
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 4 * n))]
        rp, adj = bf.csr(n, edges)
        nbrs = [set(adj[rp[u]:rp[u + 1]]) for u in range(n)]
        rpa = np.array(rp, dtype=np.uint64)
        ada = np.array(adj if adj else [0], dtype=np.int32)
        for u in range(n):
            for v in range(n):
                got = ob.oracle().orc_common(rpa.ctypes.data, ada.ctypes.data, u, v)
                assert got == len(nbrs[u] & nbrs[v])


def _check_case(n, edges, eps, mu):
    rp, adj = bf.csr(n, edges)
    labels, core, sim, st = ob.orc_scan(rp, adj, eps, mu)
    exp_labels, exp_core, exp_sim = bf.scan(n, edges, eps, mu)
    assert list(labels) == exp_labels, (n, edges, eps, mu)
    assert list(core) == [int(c) for c in exp_core]
    assert list(sim) == bf.sim_per_entry(rp, adj, exp_sim)
    return st


def test_known_answer_graphs():
    data = json.load(open(os.path.join(GOLDEN, "kat_graphs.json")))
    for case in data["cases"]:
        n, edges = case["n"], [tuple(e) for e in case["edges"]]
        rp, adj = bf.csr(n, edges)
        labels, core, sim, st = ob.orc_scan(rp, adj, case["eps"], case["mu"])
        assert list(labels) == case["labels"], case["name"]
        assert list(core) == case["core"], case["name"]
        sim_set = {tuple(e) for e in case["similar"]}
        exp_sim = {(min(u, v), max(u, v)): (min(u, v), max(u, v)) in sim_set for u, v in edges}
        assert list(sim) == bf.sim_per_entry(rp, adj, exp_sim), case["name"]
        assert st["cores"] == case["core_number"], case["name"]
        assert st["clusters"] == case["cluster_number"], case["name"]


def test_all_graphs_up_to_five_vertices():
    params = [(0.3, 1), (0.5, 2), (0.7, 2), (0.6, 3), (0.9, 0), (1.0, 1)]
    for n in range(1, 6):
        pairs = list(itertools.combinations(range(n), 2))
        for mask in range(1 << len(pairs)):
            edges = [pairs[i] for i in range(len(pairs)) if mask >> i & 1]
            for eps, mu in params:
                _check_case(n, edges, eps, mu)


def test_random_graphs_against_bruteforce():
    rng = random.Random(11)
    for trial in range(120):
        n = rng.randint(6, 48)
        p = rng.choice([0.05, 0.15, 0.3, 0.6])
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
        eps = rng.choice([0.1, 0.2, 0.35, 0.5, 0.65, 0.8])
        mu = rng.choice([-1, 0, 1, 2, 3, 5])
        _check_case(n, edges, eps, mu)


def test_similarity_entries_match_scan():
    rng = random.Random(3)
    n = 300
    edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(3000)]
    rp, adj = bf.csr(n, edges)
    for eps in (0.2, 0.5, 0.8):
        _, _, sim, _ = ob.orc_scan(rp, adj, eps, 2)
        entries = np.arange(len(adj), dtype=np.uint64)
        y, out = ob.orc_similarity_entries(rp, adj, eps, entries, threads=2)
        assert np.array_equal(out, sim)
        assert y == int(sim.sum())


def test_labels_from_sim_matches_scan():
    """orc_labels_from_sim (the phases after similarity, over given flags)
    against orc_scan on the flags orc_scan itself computes: random graphs,
    every mu convention, and R-MAT."""
    from label_check import labels_from_flags

    rng = random.Random(5)
    cases = []
    for trial in range(80):
        n = rng.randint(1, 120)
        p = rng.choice([0.02, 0.1, 0.3, 0.6])
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
        cases.append((bf.csr(n, edges), rng.choice([0.1, 0.3, 0.5, 0.7, 0.9]), rng.choice([-1, 0, 1, 2, 3, 5])))
    g = sgc.HostGraph.rmat(13, 16, seed=9)
    for eps, mu in [(0.3, 2), (0.5, 5), (0.2, 10), (0.6, 0)]:
        cases.append(((np.array(g.row_ptr), np.array(g.adj)), eps, mu))
    for (rp, adj), eps, mu in cases:
        labels, core, sim, st = ob.orc_scan(rp, adj, eps, mu)
        l2, c2, st2 = ob.orc_labels_from_sim(rp, adj, sim, mu)
        assert np.array_equal(l2, labels) and np.array_equal(c2, core), (eps, mu)
        for k in ("cores", "clusters", "borders", "hubs", "outliers", "similar_edges"):
            assert st2[k] == st[k], (k, eps, mu)


def test_labels_from_sim_on_arbitrary_flags():
    """Flags no eps produces (random symmetric subsets): orc_labels_from_sim
    against the independent vectorised rules of label_check."""
    from label_check import labels_from_flags

    g = sgc.HostGraph.rmat(12, 16, seed=3)
    rp, adj = np.array(g.row_ptr), np.array(g.adj)
    n = len(rp) - 1
    src = np.repeat(np.arange(n), np.diff(rp).astype(np.int64))
    lo, hi = np.minimum(src, adj), np.maximum(src, adj)
    for seed, frac, mu in [(1, 0.3, 2), (2, 0.6, 4), (3, 0.1, 1), (4, 0.5, 0)]:
        key = lo.astype(np.int64) * n + hi  # the same coin for both entries of an edge
        coin = (np.random.default_rng(seed).random(int(key.max()) + 1) < frac) if len(key) else np.zeros(0, bool)
        sim = coin[key].astype(np.uint8)
        labels, core, _ = ob.orc_labels_from_sim(rp, adj, sim, mu)
        l2, c2 = labels_from_flags(rp, adj, sim, mu)
        assert np.array_equal(core, c2) and np.array_equal(labels, l2), (seed, mu)


def test_prep_counts_match_scan():
    rng = random.Random(8)
    for trial in range(20):
        n = rng.randint(2, 200)
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 2000))]
        rp, adj = bf.csr(n, edges)
        for eps in (0.1, 0.3, 0.5, 0.9):
            _, _, _, st = ob.orc_scan(rp, adj, eps, 2)
            assert ob.orc_prep_counts(rp, adj, eps) == (st["prep_n"], st["prep_y"], st["prep_u"])
