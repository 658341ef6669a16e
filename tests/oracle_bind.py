"""ctypes bindings for the oracle (oracle/liboracle.so) and for the
reference's own compiled host code (oracle/_ref/libref_graph.so).

Test infrastructure: the oracle is the checker, never the thing measured.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libref_graph.so")

_VP = C.c_void_p
_orc = None


class OrcStats(C.Structure):
    _fields_ = [(name, C.c_uint64) for name in (
        "edges", "prep_n", "prep_y", "prep_u", "check_core_computed", "similar_edges",
        "cores_after_prep", "cores", "clusters", "borders", "hubs", "outliers")]

    def as_dict(self):
        return {f[0]: int(getattr(self, f[0])) for f in self._fields_}


def oracle():
    global _orc
    if _orc is not None:
        return _orc
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True, capture_output=True)
    lib = C.CDLL(ORACLE_LIB)
    lib.orc_prep.restype = C.c_int
    lib.orc_prep.argtypes = [C.c_uint32, C.c_uint32, C.c_double]
    lib.orc_predicate.restype = C.c_int
    lib.orc_predicate.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double]
    lib.orc_common.restype = C.c_uint32
    lib.orc_common.argtypes = [_VP, _VP, C.c_uint32, C.c_uint32]
    lib.orc_orient.restype = None
    lib.orc_orient.argtypes = [C.c_uint64, _VP, _VP, _VP, _VP, _VP]
    lib.orc_scan.restype = C.c_int
    lib.orc_scan.argtypes = [C.c_uint64, _VP, _VP, C.c_double, C.c_int32, _VP, _VP, _VP, C.POINTER(OrcStats)]
    lib.orc_similarity_entries.restype = C.c_uint64
    lib.orc_similarity_entries.argtypes = [C.c_uint64, _VP, _VP, C.c_double, _VP, C.c_uint64, _VP, C.c_int]
    lib.orc_labels_from_sim.restype = C.c_int
    lib.orc_labels_from_sim.argtypes = [C.c_uint64, _VP, _VP, _VP, C.c_int32, _VP, _VP, C.POINTER(OrcStats)]
    lib.orc_prep_counts.restype = None
    lib.orc_prep_counts.argtypes = [C.c_uint64, _VP, _VP, C.c_double, C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    _orc = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_VP)


def orc_scan(row_ptr, adj, eps: float, mu: int):
    """Oracle SCAN: (labels int32[n], core uint8[n], sim uint8[m], stats dict)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    ad = np.ascontiguousarray(adj, dtype=np.int32)
    n = len(rp) - 1
    m = len(ad)
    labels = np.empty(max(n, 1), dtype=np.int32)
    core = np.empty(max(n, 1), dtype=np.uint8)
    sim = np.empty(max(m, 1), dtype=np.uint8)
    st = OrcStats()
    rc = oracle().orc_scan(n, _p(rp), _p(ad if m else np.zeros(1, np.int32)), float(eps), int(mu), _p(labels),
                           _p(core), _p(sim), C.byref(st))
    assert rc == 0
    return labels[:n], core[:n], sim[:m], st.as_dict()


def orc_orient(row_ptr, adj):
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    ad = np.ascontiguousarray(adj, dtype=np.int32)
    n, m = len(rp) - 1, len(ad)
    eu = np.empty(max(m // 2, 1), dtype=np.int32)
    ev = np.empty(max(m // 2, 1), dtype=np.int32)
    eid = np.empty(max(m, 1), dtype=np.uint32)
    oracle().orc_orient(n, _p(rp), _p(ad), _p(eu), _p(ev), _p(eid))
    return eu[: m // 2], ev[: m // 2], eid[:m]


def orc_similarity_entries(row_ptr, adj, eps, entries, threads=0):
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    ad = np.ascontiguousarray(adj, dtype=np.int32)
    en = np.ascontiguousarray(entries, dtype=np.uint64)
    out = np.empty(max(len(en), 1), dtype=np.uint8)
    y = oracle().orc_similarity_entries(len(rp) - 1, _p(rp), _p(ad), float(eps), _p(en), len(en), _p(out),
                                        int(threads))
    return int(y), out[: len(en)]


def orc_labels_from_sim(row_ptr, adj, sim, mu: int):
    """The oracle's phases after similarity, over a given flag per directed
    entry: (labels int32[n], core uint8[n], stats dict)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    ad = np.ascontiguousarray(adj, dtype=np.int32)
    sm = np.ascontiguousarray(sim, dtype=np.uint8)
    n, m = len(rp) - 1, len(ad)
    assert len(sm) == m
    labels = np.empty(max(n, 1), dtype=np.int32)
    core = np.empty(max(n, 1), dtype=np.uint8)
    st = OrcStats()
    z = np.zeros(1, np.int32)
    rc = oracle().orc_labels_from_sim(n, _p(rp), _p(ad if m else z), _p(sm if m else z.view(np.uint8)), int(mu),
                                      _p(labels), _p(core), C.byref(st))
    assert rc == 0
    return labels[:n], core[:n], st.as_dict()


def orc_prep_counts(row_ptr, adj, eps: float):
    """(N, Y, U) counts of the degree-only pre-decision over every edge."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    ad = np.ascontiguousarray(adj, dtype=np.int32)
    a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
    oracle().orc_prep_counts(len(rp) - 1, _p(rp), _p(ad if len(ad) else np.zeros(1, np.int32)), float(eps),
                             C.byref(a), C.byref(b), C.byref(c))
    return int(a.value), int(b.value), int(c.value)


# ---------------------------------------------------------------------------
# The reference's own host code, built from /root/reference by
# `make -C oracle ref` (present only where the reference tree is).
# ---------------------------------------------------------------------------

_ref = None


def reference_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is not None:
        return _ref
    lib = C.CDLL(REF_LIB)
    lib.ref_load.restype = C.c_int
    lib.ref_load.argtypes = [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                             C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.POINTER(C.c_int32))]
    lib.ref_orient.restype = C.c_int
    lib.ref_orient.argtypes = [C.c_long, C.c_long, _VP, _VP, _VP, _VP, _VP]
    lib.ref_free.restype = None
    lib.ref_free.argtypes = [_VP]
    _ref = lib
    return lib


def ref_load(folder: str):
    """Graph::Graph(folder) of the reference: (row_ptr uint32[n+1], adj int32[m])."""
    lib = ref()
    n = C.c_uint64()
    m = C.c_uint64()
    rp = C.POINTER(C.c_uint32)()
    ad = C.POINTER(C.c_int32)()
    lib.ref_load(os.fsencode(folder), C.byref(n), C.byref(m), C.byref(rp), C.byref(ad))
    row_ptr = np.ctypeslib.as_array(rp, shape=(n.value + 1,)).copy()
    adj = np.ctypeslib.as_array(ad, shape=(max(m.value, 1),))[: m.value].copy()
    lib.ref_free(C.cast(rp, _VP))
    lib.ref_free(C.cast(ad, _VP))
    return row_ptr, adj


def ref_orient(row_ptr, adj):
    """getEidAndEdgeList of the reference on a CSR."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    ad = np.ascontiguousarray(adj, dtype=np.int32)
    n, m = len(rp) - 1, len(ad)
    eu = np.empty(max(m // 2, 1), dtype=np.int32)
    ev = np.empty(max(m // 2, 1), dtype=np.int32)
    eid = np.empty(max(m, 1), dtype=np.uint32)
    ref().ref_orient(n, m, _p(ad), _p(rp), _p(eu), _p(ev), _p(eid))
    return eu[: m // 2], ev[: m // 2], eid[:m]
