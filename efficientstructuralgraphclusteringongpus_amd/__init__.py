"""MI355X-native SCAN structural graph clustering (GPUSCAN++ drop-in).

Python host mirror over the C ABI of ``lib/libsgc.so`` (declared in
``include/sgc.h``).  The compute path is the HIP library; this module only
binds it with ctypes and moves numpy arrays.  There is no CPU fallback: if
the library is missing or cannot be loaded every call raises.

The reference's surfaces mirrored here:

* ``GPUScan(folder, eps, miu)``: the ``GPUScan folder_path epsilon miu`` run
  (GPUSCAN++/main.cpp:20-123) returning the per-vertex labels that
  ``int* GPUScan(graph_t*, ...)`` (GPUSCAN++/GPUScan.h:7) returns.
* ``HostGraph.load(dir)``: ``Graph::Graph(dir)`` (GPUSCAN++/Graph.cpp:147-153).
* ``HostGraph.orient()``: ``getEidAndEdgeList`` (GPUSCAN++/Graph.cpp:52-141).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, fields

import numpy as np

__all__ = [
    "SGCError", "HostGraph", "Context", "DeviceGraph", "ScanStats", "ScanResult",
    "scan", "scan_device", "scan_ooc", "ooc_min_budget", "OocStats", "Server", "QueryStats", "Tuning",
    "copy_bench", "scan_device_noalloc", "noalloc_max_bytes", "noalloc_bytes", "write_result", "read_result", "ROLE_NAMES", "GPUScan", "library_path", "load_library", "HUB", "OUTLIER",
]

HUB = -1
OUTLIER = -2

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def library_path() -> str:
    """lib/libsgc.so of this package, or the build named by SGC_LIBRARY."""
    return os.environ.get("SGC_LIBRARY") or os.path.join(_PKG, "lib", "libsgc.so")


class SGCError(RuntimeError):
    """A libsgc call failed; ``code`` is the SGC_E* status."""

    CODES = {1: "EINVAL", 2: "EIO", 3: "EFORMAT", 4: "ENOMEM", 5: "EDEVICE", 6: "ERANGE", 7: "EBUSY"}

    def __init__(self, code: int, message: str):
        super().__init__(f"[{self.CODES.get(code, code)}] {message}")
        self.code = code
        self.message = message


class _Stats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double), ("ms_prep", C.c_double), ("ms_similarity", C.c_double),
        ("ms_core", C.c_double), ("ms_cluster", C.c_double), ("ms_border", C.c_double),
        ("ms_classify", C.c_double), ("ms_similarity_large", C.c_double),
        ("edges", C.c_uint64), ("u_edges", C.c_uint64), ("similar_edges", C.c_uint64),
        ("cores", C.c_uint64), ("clusters", C.c_uint64), ("borders", C.c_uint64),
        ("hubs", C.c_uint64), ("outliers", C.c_uint64),
        ("alg_bytes", C.c_double), ("probe_elements", C.c_double),
        ("loaded_elements", C.c_double), ("claims", C.c_double),
        ("deferred_edges", C.c_uint64), ("redo_edges", C.c_uint64), ("pruned", C.c_uint32),
    ]


class _OocStats(C.Structure):
    _fields_ = [
        ("device_bytes", C.c_uint64), ("similarity_chips", C.c_uint32), ("row_chips", C.c_uint32),
        ("bytes_h2d", C.c_double), ("bytes_d2h", C.c_double), ("ms_host", C.c_double), ("ms_phases", C.c_double),
        ("row_blocks", C.c_uint32), ("blocks_per_group", C.c_uint32),
    ]


@dataclass
class OocStats:
    device_bytes: int
    similarity_chips: int
    row_chips: int
    bytes_h2d: float
    bytes_d2h: float
    ms_host: float
    ms_phases: float
    row_blocks: int
    blocks_per_group: int


class _QueryStats(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("graph", C.c_uint32), ("n", C.c_uint64), ("ms_device", C.c_double),
        ("ms_latency", C.c_double), ("similar_edges", C.c_uint64), ("cores", C.c_uint64), ("clusters", C.c_uint64),
        ("borders", C.c_uint64), ("hubs", C.c_uint64), ("outliers", C.c_uint64),
    ]


@dataclass
class QueryStats:
    status: int
    graph: int
    n: int
    ms_device: float
    ms_latency: float
    similar_edges: int
    cores: int
    clusters: int
    borders: int
    hubs: int
    outliers: int


class _Tuning(C.Structure):
    _fields_ = [("small_max_degree", C.c_uint32), ("small_unit_work", C.c_double), ("large_unit_work", C.c_double),
                ("hash_tables", C.c_uint32), ("tiny_work", C.c_uint32), ("ms_similarity", C.c_double),
                ("ms_default", C.c_double), ("trials", C.c_uint32)]


@dataclass
class Tuning:
    small_max_degree: int = 1024
    small_unit_work: float = 262144.0
    large_unit_work: float = 8388608.0
    hash_tables: int = 1
    tiny_work: int = 2048
    ms_similarity: float = 0.0
    ms_default: float = 0.0
    trials: int = 0


class _GraphT(C.Structure):
    _fields_ = [("n", C.c_long), ("m", C.c_long), ("adj", C.POINTER(C.c_int32)),
                ("num_edges", C.POINTER(C.c_uint32)), ("eid", C.POINTER(C.c_uint32))]


@dataclass
class ScanStats:
    ms_total: float
    ms_prep: float
    ms_similarity: float
    ms_core: float
    ms_cluster: float
    ms_border: float
    ms_classify: float
    ms_similarity_large: float
    edges: int
    u_edges: int
    similar_edges: int
    cores: int
    clusters: int
    borders: int
    hubs: int
    outliers: int
    alg_bytes: float
    probe_elements: float
    loaded_elements: float
    claims: float
    deferred_edges: int = 0
    redo_edges: int = 0
    pruned: int = 0

    @classmethod
    def _from(cls, s: _Stats) -> "ScanStats":
        vals = {f.name: getattr(s, f.name) for f in fields(cls)}
        return cls(**vals)


@dataclass
class ScanResult:
    labels: np.ndarray          # int32[n]: cluster id (min core id), -1 hub, -2 outlier
    core: np.ndarray            # uint8[n]
    sim: np.ndarray | None      # uint8[m] per directed CSR entry (1 = similar)
    stats: ScanStats


_VP = C.c_void_p
_P = C.POINTER


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def load_library(path: str | None = None):
    """Loads libsgc.so once; raises if it is missing (no fallback path)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or library_path()
    if not os.path.exists(path):
        raise SGCError(5, f"{path} is not built; run efficientstructuralgraphclusteringongpus_amd.build")
    lib = C.CDLL(path)
    _sig(lib, "sgc_last_error", C.c_char_p, [])
    _sig(lib, "sgc_version", C.c_char_p, [])
    _sig(lib, "sgc_graph_load", C.c_int, [C.c_char_p, C.c_int, _P(_VP)])
    _sig(lib, "sgc_graph_save", C.c_int, [_VP, C.c_char_p])
    _sig(lib, "sgc_graph_from_csr", C.c_int, [C.c_uint64, C.c_uint64, _VP, _VP, C.c_int, _P(_VP)])
    _sig(lib, "sgc_graph_from_edges", C.c_int, [C.c_uint64, C.c_uint64, _VP, _VP, _P(_VP)])
    _sig(lib, "sgc_graph_convert_edge_list", C.c_int, [C.c_char_p, C.c_char_p])
    _sig(lib, "sgc_graph_rmat", C.c_int,
         [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, _P(_VP)])
    _sig(lib, "sgc_graph_planted", C.c_int,
         [C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_uint32,
          C.c_double, C.c_uint64, _P(_VP)])
    _sig(lib, "sgc_graph_n", C.c_uint64, [_VP])
    _sig(lib, "sgc_graph_m", C.c_uint64, [_VP])
    _sig(lib, "sgc_graph_row_ptr", _P(C.c_uint64), [_VP])
    _sig(lib, "sgc_graph_adj", _P(C.c_int32), [_VP])
    _sig(lib, "sgc_graph_max_degree", C.c_uint32, [_VP])
    _sig(lib, "sgc_graph_digest", C.c_uint64, [_VP])
    _sig(lib, "sgc_graph_free", None, [_VP])
    _sig(lib, "sgc_graph_pin", C.c_int, [_VP])
    _sig(lib, "sgc_graph_pinned", C.c_int, [_VP])
    _sig(lib, "sgc_graph_similarity_work", C.c_int,
         [_VP, C.c_double, _P(C.c_uint64), _P(C.c_double), _P(C.c_double), _P(C.c_double)])
    _sig(lib, "sgc_philox4x32_10", None, [_VP, _VP, _VP])
    _sig(lib, "sgc_orient", C.c_int, [_VP, _VP, _VP, _VP])
    _sig(lib, "sgc_ctx_create", C.c_int, [C.c_int, _P(_VP)])
    _sig(lib, "sgc_ctx_destroy", None, [_VP])
    _sig(lib, "sgc_ctx_sync", C.c_int, [_VP])
    _sig(lib, "sgc_ctx_device_name", C.c_char_p, [_VP])
    _sig(lib, "sgc_ctx_memory", C.c_int, [_VP, _P(C.c_uint64), _P(C.c_uint64)])
    _sig(lib, "sgc_ctx_set_similarity_kernel", C.c_int, [_VP, C.c_int])
    _sig(lib, "sgc_ctx_set_pruning", C.c_int, [_VP, C.c_int])
    _sig(lib, "sgc_dgraph_upload", C.c_int, [_VP, _VP, _P(_VP)])
    _sig(lib, "sgc_dgraph_free", None, [_VP])
    _sig(lib, "sgc_dgraph_n", C.c_uint64, [_VP])
    _sig(lib, "sgc_dgraph_m", C.c_uint64, [_VP])
    _sig(lib, "sgc_scan", C.c_int, [_VP, _VP, C.c_double, C.c_int32, _VP, _VP, _VP, _P(_Stats)])
    _sig(lib, "sgc_ooc_min_budget", C.c_int, [_VP, C.c_double, _P(C.c_uint64)])
    _sig(lib, "sgc_write_result", C.c_int,
         [C.c_char_p, C.c_uint64, _VP, _VP, C.c_double, C.c_int32, _P(_Stats), C.c_int])
    _sig(lib, "sgc_dgraph_get_tuning", C.c_int, [_VP, _P(_Tuning)])
    _sig(lib, "sgc_dgraph_set_tuning", C.c_int, [_VP, _P(_Tuning)])
    _sig(lib, "sgc_dgraph_autotune", C.c_int, [_VP, _VP, C.c_double, _P(_Tuning)])
    _sig(lib, "sgc_server_create", C.c_int, [_VP, C.c_uint64, C.c_uint32, _P(_VP)])
    _sig(lib, "sgc_server_destroy", None, [_VP])
    _sig(lib, "sgc_server_graph_bytes", C.c_uint64, [_VP, C.c_uint32])
    _sig(lib, "sgc_server_bytes_in_use", C.c_uint64, [_VP])
    _sig(lib, "sgc_server_set_graph_launch", C.c_int, [_VP, C.c_int])
    _sig(lib, "sgc_server_set_batching", C.c_int, [_VP, C.c_uint64, C.c_uint32])
    _sig(lib, "sgc_server_get_batching", C.c_int, [_VP, _P(C.c_uint64), _P(C.c_uint32)])
    _sig(lib, "sgc_server_autotune_batching", C.c_int, [_VP, C.c_double, C.c_int32, _P(C.c_uint64)])
    _sig(lib, "sgc_server_add_graph", C.c_int, [_VP, _VP, _P(C.c_uint32), _P(C.c_double)])
    _sig(lib, "sgc_server_remove_graph", C.c_int, [_VP, C.c_uint32])
    _sig(lib, "sgc_server_submit", C.c_int, [_VP, C.c_uint32, C.c_double, C.c_int32, _P(C.c_uint64)])
    _sig(lib, "sgc_server_progress", C.c_int, [_VP, C.c_int, _P(C.c_uint32)])
    _sig(lib, "sgc_server_pending", C.c_uint32, [_VP])
    _sig(lib, "sgc_server_completed", C.c_int, [_VP, _VP, C.c_uint32, _P(C.c_uint32)])
    _sig(lib, "sgc_server_result", C.c_int, [_VP, C.c_uint64, _VP, C.c_uint64, _P(_QueryStats)])
    _sig(lib, "sgc_scan_ooc", C.c_int,
         [_VP, _VP, C.c_double, C.c_int32, C.c_uint64, _VP, _VP, _VP, _P(_Stats), _P(_OocStats)])
    _sig(lib, "sgc_scan_device", C.c_int, [_VP, _VP, C.c_double, C.c_int32, _VP, _VP, _VP, _P(_Stats)])
    _sig(lib, "sgc_scan_device_noalloc", C.c_int, [_VP, _VP, C.c_double, C.c_int32, _VP, _VP, _VP, _P(_Stats)])
    _sig(lib, "sgc_noalloc_max_bytes", C.c_uint64, [])
    _sig(lib, "sgc_noalloc_bytes", C.c_uint64, [_VP])
    _sig(lib, "sgc_GPUScan", _P(C.c_int32), [_P(_GraphT), _VP, _VP, C.c_double, C.c_int32])
    _sig(lib, "sgc_copy_bench", C.c_int, [_VP, C.c_uint64, _P(C.c_double)])
    _LIB = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise SGCError(rc, _LIB.sgc_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_VP)


class HostGraph:
    """Host CSR graph owned by libsgc (``sgc_host_graph``)."""

    def __init__(self, handle):
        self._h = handle
        self._lib = load_library()

    def __del__(self):
        self.close()

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sgc_graph_free(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @classmethod
    def load(cls, folder: str, check_symmetry: bool = False) -> "HostGraph":
        lib = load_library()
        h = _VP()
        _check(lib.sgc_graph_load(os.fsencode(folder), int(check_symmetry), C.byref(h)))
        return cls(h)

    @classmethod
    def from_csr(cls, row_ptr, adj, check_symmetry: bool = True) -> "HostGraph":
        lib = load_library()
        rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        ad = np.ascontiguousarray(adj, dtype=np.int32)
        h = _VP()
        _check(lib.sgc_graph_from_csr(len(rp) - 1, len(ad), _ptr(rp), _ptr(ad), int(check_symmetry), C.byref(h)))
        return cls(h)

    @classmethod
    def from_edges(cls, n: int, src, dst) -> "HostGraph":
        lib = load_library()
        s = np.ascontiguousarray(src, dtype=np.int32)
        d = np.ascontiguousarray(dst, dtype=np.int32)
        h = _VP()
        _check(lib.sgc_graph_from_edges(n, len(s), _ptr(s), _ptr(d), C.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, scale: int, edge_factor: float = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1) -> "HostGraph":
        lib = load_library()
        h = _VP()
        _check(lib.sgc_graph_rmat(scale, edge_factor, a, b, c, seed, C.byref(h)))
        return cls(h)

    @classmethod
    def planted(cls, n: int = 2_000_000, communities: int = 20_000, zipf: float = 1.5, p_in: float = 0.05,
                max_intra_degree: float = 200.0, inter_per_vertex: int = 8, hubs: int = 64,
                hub_fraction: float = 0.05, seed: int = 1) -> "HostGraph":
        lib = load_library()
        h = _VP()
        _check(lib.sgc_graph_planted(n, communities, zipf, p_in, max_intra_degree, inter_per_vertex, hubs,
                                     hub_fraction, seed, C.byref(h)))
        return cls(h)

    @staticmethod
    def convert_edge_list(txt_path: str, out_dir: str) -> None:
        lib = load_library()
        _check(lib.sgc_graph_convert_edge_list(os.fsencode(txt_path), os.fsencode(out_dir)))

    def save(self, folder: str) -> None:
        os.makedirs(folder, exist_ok=True)
        _check(self._lib.sgc_graph_save(self._h, os.fsencode(folder)))

    @property
    def n(self) -> int:
        return int(self._lib.sgc_graph_n(self._h))

    @property
    def m(self) -> int:
        return int(self._lib.sgc_graph_m(self._h))

    @property
    def row_ptr(self) -> np.ndarray:
        """Read-only view (valid while the graph is alive)."""
        p = self._lib.sgc_graph_row_ptr(self._h)
        a = np.ctypeslib.as_array(p, shape=(self.n + 1,))
        a.flags.writeable = False
        return a

    @property
    def adj(self) -> np.ndarray:
        m = self.m
        if m == 0:
            return np.zeros(0, dtype=np.int32)
        a = np.ctypeslib.as_array(self._lib.sgc_graph_adj(self._h), shape=(m,))
        a.flags.writeable = False
        return a

    @property
    def max_degree(self) -> int:
        return int(self._lib.sgc_graph_max_degree(self._h))

    def pin(self) -> None:
        """Moves the adjacency into page-locked host memory (needs a device);
        scan_ooc then sends row blocks without a staging copy."""
        _check(self._lib.sgc_graph_pin(self._h))

    @property
    def pinned(self) -> bool:
        return bool(self._lib.sgc_graph_pinned(self._h))

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptr).astype(np.int64)

    def digest(self) -> int:
        return int(self._lib.sgc_graph_digest(self._h))

    def similarity_work(self, eps: float):
        """(undecided edges, algorithmic bytes, short-row elements, algorithmic
        bytes of the edges whose owner has degree > 1024) at eps."""
        u = C.c_uint64()
        b = C.c_double()
        sh = C.c_double()
        bl = C.c_double()
        _check(self._lib.sgc_graph_similarity_work(self._h, float(eps), C.byref(u), C.byref(b), C.byref(sh),
                                                   C.byref(bl)))
        return int(u.value), float(b.value), float(sh.value), float(bl.value)

    def orient(self):
        """(edge_u, edge_v, eid) as getEidAndEdgeList builds them."""
        E = self.m // 2
        eu = np.empty(max(E, 1), dtype=np.int32)
        ev = np.empty(max(E, 1), dtype=np.int32)
        eid = np.empty(max(self.m, 1), dtype=np.uint32)
        _check(self._lib.sgc_orient(self._h, _ptr(eu), _ptr(ev), _ptr(eid)))
        return eu[:E], ev[:E], eid[: self.m]


class Context:
    """A HIP device and stream (``sgc_ctx``)."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = _VP()
        _check(self._lib.sgc_ctx_create(device, C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sgc_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def device_name(self) -> str:
        return self._lib.sgc_ctx_device_name(self._h).decode()

    def sync(self) -> None:
        _check(self._lib.sgc_ctx_sync(self._h))

    def memory(self) -> tuple[int, int]:
        """(free, total) bytes of device memory."""
        f, t = C.c_uint64(), C.c_uint64()
        _check(self._lib.sgc_ctx_memory(self._h, C.byref(f), C.byref(t)))
        return f.value, t.value

    SIM_TIERED = 0
    SIM_WAVE_PER_EDGE = 1

    def set_similarity_kernel(self, kind: int) -> None:
        _check(self._lib.sgc_ctx_set_similarity_kernel(self._h, int(kind)))

    def set_pruning(self, enable: bool) -> None:
        """Core-check pruning in passes that return no similarity flags
        (default off; results are the same either way)."""
        _check(self._lib.sgc_ctx_set_pruning(self._h, int(bool(enable))))


class DeviceGraph:
    """A CSR graph resident in HBM (``sgc_dgraph``)."""

    def __init__(self, ctx: Context, graph: HostGraph):
        self._lib = load_library()
        self.ctx = ctx
        h = _VP()
        _check(self._lib.sgc_dgraph_upload(ctx.handle, graph.handle, C.byref(h)))
        self._h = h
        self.n = graph.n
        self.m = graph.m

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sgc_dgraph_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def tuning(self) -> Tuning:
        t = _Tuning()
        _check(self._lib.sgc_dgraph_get_tuning(self._h, C.byref(t)))
        return Tuning(**{f.name: getattr(t, f.name) for f in fields(Tuning)})

    @tuning.setter
    def tuning(self, value: Tuning) -> None:
        t = _Tuning(value.small_max_degree, value.small_unit_work, value.large_unit_work, value.hash_tables,
                    value.tiny_work, 0.0, 0.0, 0)
        _check(self._lib.sgc_dgraph_set_tuning(self._h, C.byref(t)))

    def autotune(self, eps: float) -> Tuning:
        """Times the similarity phase at eps under candidate planner knobs and
        keeps the fastest on this graph."""
        t = _Tuning()
        _check(self._lib.sgc_dgraph_autotune(self.ctx.handle, self._h, float(eps), C.byref(t)))
        return Tuning(**{f.name: getattr(t, f.name) for f in fields(Tuning)})


def scan(ctx: Context, dg: DeviceGraph, eps: float, mu: int, want_sim: bool = False) -> ScanResult:
    """One plain SCAN pass on a resident graph; returns host copies."""
    lib = load_library()
    labels = np.empty(max(dg.n, 1), dtype=np.int32)
    core = np.empty(max(dg.n, 1), dtype=np.uint8)
    sim = np.empty(max(dg.m, 1), dtype=np.uint8) if want_sim else None
    st = _Stats()
    _check(lib.sgc_scan(ctx.handle, dg.handle, float(eps), int(mu), _ptr(labels), _ptr(core), _ptr(sim),
                        C.byref(st)))
    return ScanResult(labels[: dg.n], core[: dg.n], None if sim is None else sim[: dg.m], ScanStats._from(st))


def ooc_min_budget(g: HostGraph, eps: float) -> int:
    """Smallest device budget scan_ooc accepts for g at eps (bytes)."""
    lib = load_library()
    b = C.c_uint64()
    _check(lib.sgc_ooc_min_budget(g._h, float(eps), C.byref(b)))
    return int(b.value)


def scan_ooc(ctx: Context, g: HostGraph, eps: float, mu: int, device_budget: int,
             want_sim: bool = False) -> tuple[ScanResult, OocStats]:
    """Out-of-core pass over a host-resident graph of any size within
    `device_budget` bytes of device memory; results identical to scan()."""
    lib = load_library()
    n, m = g.n, g.m
    labels = np.empty(max(n, 1), dtype=np.int32)
    core = np.empty(max(n, 1), dtype=np.uint8)
    sim = np.empty(max(m, 1), dtype=np.uint8) if want_sim else None
    st = _Stats()
    oi = _OocStats()
    _check(lib.sgc_scan_ooc(ctx.handle, g._h, float(eps), int(mu), int(device_budget), _ptr(labels), _ptr(core),
                            _ptr(sim), C.byref(st), C.byref(oi)))
    info = OocStats(**{f.name: getattr(oi, f.name) for f in fields(OocStats)})
    return ScanResult(labels[:n], core[:n], None if sim is None else sim[:m], ScanStats._from(st)), info


def scan_device(ctx: Context, dg: DeviceGraph, eps: float, mu: int, d_labels: int | None = None,
                d_core: int | None = None, d_sim: int | None = None) -> ScanStats:
    """One pass leaving results in HBM (optionally copied to caller device
    buffers given as raw pointers); returns the pass statistics."""
    lib = load_library()
    st = _Stats()
    _check(lib.sgc_scan_device(ctx.handle, dg.handle, float(eps), int(mu), d_labels, d_core, d_sim,
                               C.byref(st)))
    return ScanStats._from(st)


def scan_device_noalloc(ctx: Context, dg: DeviceGraph, eps: float, mu: int, d_labels: int | None = None,
                        d_core: int | None = None, d_sim: int | None = None) -> ScanStats:
    """scan_device without any device allocation: the pass's work arrays live
    in the library's static per-GPU scratch (noalloc_max_bytes() bytes);
    raises SGCError (ERANGE) for a graph whose pass needs more."""
    lib = load_library()
    st = _Stats()
    _check(lib.sgc_scan_device_noalloc(ctx.handle, dg.handle, float(eps), int(mu), d_labels, d_core, d_sim,
                                       C.byref(st)))
    return ScanStats._from(st)


def noalloc_max_bytes() -> int:
    return int(load_library().sgc_noalloc_max_bytes())


def noalloc_bytes(dg: DeviceGraph) -> int:
    return int(load_library().sgc_noalloc_bytes(dg.handle))


def copy_bench(ctx: Context, nbytes: int) -> float:
    """Milliseconds of one plain device copy of nbytes (HBM reference point)."""
    lib = load_library()
    ms = C.c_double()
    _check(lib.sgc_copy_bench(ctx.handle, int(nbytes), C.byref(ms)))
    return ms.value


class Server:
    """Serving engine: resident graphs with precomputed common-neighbour
    counts, many concurrent (graph, eps, mu) queries."""

    EBUSY = 7

    def __init__(self, ctx: Context, hbm_budget: int, slots_per_graph: int = 4):
        self._lib = load_library()
        self._ctx = ctx
        h = _VP()
        _check(self._lib.sgc_server_create(ctx.handle, int(hbm_budget), int(slots_per_graph), C.byref(h)))
        self._h = h.value
        self.slots_per_graph = slots_per_graph
        self._n = {}

    @staticmethod
    def graph_bytes(g: HostGraph, slots_per_graph: int = 4) -> int:
        return int(load_library().sgc_server_graph_bytes(g._h, int(slots_per_graph)))

    @property
    def bytes_in_use(self) -> int:
        return int(self._lib.sgc_server_bytes_in_use(self._h))

    def set_graph_launch(self, enable: bool) -> None:
        """Replay a captured hipGraph per query (default) or launch eagerly."""
        _check(self._lib.sgc_server_set_graph_launch(self._h, 1 if enable else 0))

    def set_batching(self, cutoff_entries: int, max_queries: int = 64) -> None:
        """Queries on graphs of at most cutoff_entries entries run in batches
        of up to max_queries, one launch per phase (0: off, the default)."""
        _check(self._lib.sgc_server_set_batching(self._h, int(cutoff_entries), int(max_queries)))

    def batching(self) -> tuple[int, int]:
        cut, mx = C.c_uint64(), C.c_uint32()
        _check(self._lib.sgc_server_get_batching(self._h, C.byref(cut), C.byref(mx)))
        return int(cut.value), int(mx.value)

    def autotune_batching(self, eps: float = 0.5, mu: int = 5) -> int:
        """Times replay against batches per admitted graph and sets the
        cut-off; returns it (entries)."""
        cut = C.c_uint64()
        _check(self._lib.sgc_server_autotune_batching(self._h, float(eps), int(mu), C.byref(cut)))
        return int(cut.value)

    def add_graph(self, g: HostGraph) -> tuple[int, float]:
        gid = C.c_uint32()
        ms = C.c_double()
        _check(self._lib.sgc_server_add_graph(self._h, g._h, C.byref(gid), C.byref(ms)))
        self._n[gid.value] = g.n
        return gid.value, ms.value

    def remove_graph(self, gid: int) -> None:
        _check(self._lib.sgc_server_remove_graph(self._h, int(gid)))

    def submit(self, gid: int, eps: float, mu: int) -> int:
        t = C.c_uint64()
        _check(self._lib.sgc_server_submit(self._h, int(gid), float(eps), int(mu), C.byref(t)))
        return t.value

    def progress(self, wait: bool = False) -> int:
        done = C.c_uint32()
        _check(self._lib.sgc_server_progress(self._h, 1 if wait else 0, C.byref(done)))
        return done.value

    @property
    def pending(self) -> int:
        return int(self._lib.sgc_server_pending(self._h))

    def completed(self, capacity: int = 4096) -> list[int]:
        """Tickets retired since the last call (each listed once)."""
        buf = np.empty(capacity, dtype=np.uint64)
        count = C.c_uint32()
        _check(self._lib.sgc_server_completed(self._h, _ptr(buf), int(capacity), C.byref(count)))
        return [int(t) for t in buf[: count.value]]

    def result(self, ticket: int, n: int):
        """(labels, QueryStats) of a retired query, or None while it runs."""
        labels = np.empty(max(n, 1), dtype=np.int32)
        st = _QueryStats()
        rc = self._lib.sgc_server_result(self._h, int(ticket), _ptr(labels), labels.size, C.byref(st))
        if rc == self.EBUSY:
            return None
        _check(rc)
        return labels[:n], QueryStats(**{f.name: getattr(st, f.name) for f in fields(QueryStats)})

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sgc_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


ROLE_NAMES = ("core", "border", "hub", "outlier")


def write_result(out_dir: str, result: ScanResult, eps: float, mu: int, text: bool = False) -> None:
    """labels.bin / core.bin / role.bin / summary.txt (and labels.txt) in out_dir."""
    lib = load_library()
    labels = np.ascontiguousarray(result.labels, dtype=np.int32)
    core = np.ascontiguousarray(result.core, dtype=np.uint8)
    st = _Stats()
    for f in fields(ScanStats):
        setattr(st, f.name, getattr(result.stats, f.name))
    _check(lib.sgc_write_result(out_dir.encode(), len(labels), _ptr(labels), _ptr(core), float(eps), int(mu),
                                C.byref(st), 1 if text else 0))


def read_result(out_dir: str) -> dict:
    """The arrays and summary fields sgc_write_result wrote (no pickle: raw
    little-endian arrays and "key: value" text)."""
    out = {
        "labels": np.fromfile(os.path.join(out_dir, "labels.bin"), dtype="<i4"),
        "core": np.fromfile(os.path.join(out_dir, "core.bin"), dtype=np.uint8),
        "role": np.fromfile(os.path.join(out_dir, "role.bin"), dtype=np.uint8),
    }
    summary = {}
    for line in open(os.path.join(out_dir, "summary.txt")):
        k, _, v = line.rstrip("\n").partition(": ")
        summary[k] = v
    out["summary"] = summary
    return out


def GPUScan(folder: str, eps: float, miu: int) -> np.ndarray:
    """The reference's run on ``folder`` (b_degree.bin/b_adj.bin): labels[n]."""
    g = HostGraph.load(folder)
    with Context(0) as ctx:
        dg = DeviceGraph(ctx, g)
        try:
            return scan(ctx, dg, eps, miu).labels
        finally:
            dg.close()
