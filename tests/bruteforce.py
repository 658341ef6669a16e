"""Declarative SCAN on tiny graphs, used to pin the oracle.

Written from the definitions, not from the oracle's state machine:

* similarity of an edge: the degree-only decision of GPUSCAN++/GPUScan.cu:
  143-169 first, then the live predicate of :220-221 on the common-neighbour
  count of python sets;
* core: at least mu similar neighbours (mu <= 0: at least one);
* clusters: connected components of cores over similar core-core edges,
  found by breadth-first search, labelled by their minimum vertex;
* border: a non-core with a similar core neighbour takes the minimum label
  among those neighbours' clusters;
* the rest: mu <= 0 -> hub (-1); otherwise hub iff its neighbours carry at
  least two distinct labels, else outlier (-2).
"""
from __future__ import annotations

from collections import deque


def wrap_i32(x: int) -> int:
    x &= 0xFFFFFFFF
    return x - (1 << 32) if x >= (1 << 31) else x


def edge_similar(nbrs, x: int, y: int, eps: float) -> bool:
    dx, dy = len(nbrs[x]), len(nbrs[y])
    a, b = (x, y) if (dx, x) < (dy, y) else (y, x)
    A, B = float(len(nbrs[a]) + 1), float(len(nbrs[b]) + 1)
    if A < B * eps * eps:
        return False
    if eps * eps * A * B <= 4.0:
        return True
    cn = len(nbrs[a] & nbrs[b])
    lhs = wrap_i32((cn + 2) * (cn + 2))
    rhs = eps * eps * float(len(nbrs[a]) + 1) * float(len(nbrs[b]) + 1)
    return not (float(lhs) < rhs)


def scan(n: int, edges, eps: float, mu: int):
    """Returns (labels, core, sim) with sim as a dict {(min, max): bool}."""
    nbrs = [set() for _ in range(n)]
    for u, v in edges:
        if u != v:
            nbrs[u].add(v)
            nbrs[v].add(u)
    sim = {}
    for u in range(n):
        for v in nbrs[u]:
            if u < v:
                sim[(u, v)] = edge_similar(nbrs, u, v, eps)

    def s(u, v):
        return sim[(min(u, v), max(u, v))]

    thr = mu if mu >= 1 else 1
    core = [sum(1 for v in nbrs[u] if s(u, v)) >= thr for u in range(n)]
    label = [None] * n
    for u in range(n):
        if core[u] and label[u] is None:
            comp, q = [u], deque([u])
            label[u] = -3
            while q:
                x = q.popleft()
                for y in nbrs[x]:
                    if core[y] and label[y] is None and s(x, y):
                        label[y] = -3
                        comp.append(y)
                        q.append(y)
            lab = min(comp)
            for x in comp:
                label[x] = lab
    border = [None] * n
    for x in range(n):
        if not core[x]:
            cands = [label[c] for c in nbrs[x] if core[c] and s(x, c)]
            if cands:
                border[x] = min(cands)
    out = []
    for x in range(n):
        if core[x]:
            out.append(label[x])
        elif border[x] is not None:
            out.append(border[x])
        elif mu <= 0:
            out.append(-1)
        else:
            labs = set()
            for y in nbrs[x]:
                ly = label[y] if core[y] else border[y]
                if ly is not None:
                    labs.add(ly)
            out.append(-1 if len(labs) >= 2 else -2)
    return out, core, sim


def csr(n: int, edges):
    """Symmetric, sorted, self-loop-free CSR of an edge list."""
    nbrs = [set() for _ in range(n)]
    for u, v in edges:
        if u != v:
            nbrs[u].add(v)
            nbrs[v].add(u)
    row_ptr = [0]
    adj = []
    for u in range(n):
        adj.extend(sorted(nbrs[u]))
        row_ptr.append(len(adj))
    return row_ptr, adj


def sim_per_entry(row_ptr, adj, sim):
    out = []
    for u in range(len(row_ptr) - 1):
        for k in range(row_ptr[u], row_ptr[u + 1]):
            v = adj[k]
            out.append(1 if sim[(min(u, v), max(u, v))] else 0)
    return out
