"""Vectorised restatement of the label rules given per-entry similarity flags.

Used at sizes where the sequential oracle is slow: given the similarity of
every directed entry (checked against the oracle on a sample), the cores,
clusters (minimum core id per connected component), borders (minimum
candidate) and hubs/outliers follow from the definitions alone.  Pinned
against the oracle in tests/test_label_check.py.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components


def labels_from_flags(row_ptr, adj, sim, mu: int):
    rp = np.asarray(row_ptr, dtype=np.int64)
    adj = np.asarray(adj, dtype=np.int64)
    sim = np.asarray(sim).astype(bool)
    n = len(rp) - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    sd = np.bincount(src[sim], minlength=n)
    thr = mu if mu >= 1 else 1
    core = sd >= thr
    cc = sim & core[src] & core[adj]
    g = coo_matrix((np.ones(int(cc.sum()), dtype=np.int8), (src[cc], adj[cc])), shape=(n, n))
    _, comp = connected_components(g, directed=False)
    big = np.iinfo(np.int64).max
    comp_min = np.full(n, big, dtype=np.int64)
    ids = np.nonzero(core)[0]
    np.minimum.at(comp_min, comp[ids], ids)
    label = np.full(n, big, dtype=np.int64)
    label[core] = comp_min[comp[core]]
    bm = sim & ~core[src] & core[adj]
    np.minimum.at(label, src[bm], label[adj[bm]])
    rest = label == big
    if mu <= 0:
        label[rest] = -1
    else:
        has = label != big
        em = rest[src] & has[adj]
        lo = np.full(n, big, dtype=np.int64)
        hi = np.full(n, -1, dtype=np.int64)
        np.minimum.at(lo, src[em], label[adj[em]])
        np.maximum.at(hi, src[em], label[adj[em]])
        hub = rest & (lo != big) & (lo != hi)
        label[rest] = -2
        label[hub] = -1
    return label.astype(np.int32), core.astype(np.uint8)
