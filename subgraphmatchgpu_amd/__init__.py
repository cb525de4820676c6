"""MI355X subgraph enumeration engine: Python view of the C ABI in include/sme.h.

The compute path is libsme.so (hand-written gfx950 HIP kernels behind an
extern "C" boundary).  This module only loads it with ctypes and moves numpy
arrays across; there is no Python or CPU fallback.  If the library is missing
every entry point raises.

Mirrors the reference's host interface for the GPU path:
  Query ids / QueryType              src/Meta.h:104-124
  Algo ids (GPU_GPSM=4, GPU_LIGHT=6) src/Meta.h:47-66
  compress levels                    src/query/common/LazyTraversalCommon.h:38-76
  count over a start-vertex range    HybridGPUComponent::GPUThreadExecute
                                     (src/gpu/impl/common/HybridGPUComp.cuh:53-69)
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsme.so")
CLI_PATH = os.path.join(_HERE, "patternmatchgpu")

# QueryType (src/Meta.h:104-124)
Q0, Q1, Q2, Q3, Q4, Q5, Q6, Q7, Q8, Q9, Q10, Q11, Q12, Q13 = range(14)
QUERY_NAMES = {
    0: "triangle", 1: "square", 2: "chordal_square", 3: "4clique", 4: "house",
    5: "quad_triangle", 6: "near5clique", 7: "5clique", 8: "chordal_roof",
    9: "three_triangle", 10: "solar_square", 11: "6clique", 12: "near_near_5clique",
    13: "triangle_in_diamond",
}
# Algo (src/Meta.h:47-66)
GPU_GPSM, GPU_GPSM_REUSE, GPU_LIGHT = 4, 5, 6
# LazyTraversalCompressLevel
CL_MATERIALIZE, CL_NON_MATERIALIZE, CL_NON_MATERIALIZE_OPT, CL_SPECIAL = range(4)

STATUS = {0: "OK", 1: "INVALID", 2: "OOM", 3: "HIP", 4: "IO", 5: "UNSUPPORTED"}
MAX_LEVELS = 16


class SmeError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class _Stats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_uint64), ("d2h_syncs", ctypes.c_uint64),
                ("peak_pool_bytes", ctypes.c_uint64), ("elapsed_ms", ctypes.c_double),
                ("ranked", ctypes.c_uint64)]


class _Tuning(ctypes.Structure):
    _fields_ = [("lane_path_max", ctypes.c_uint64), ("item_grid", ctypes.c_uint32),
                ("super_chunks", ctypes.c_uint32), ("batch_bytes", ctypes.c_uint64),
                ("stream_ratio", ctypes.c_uint32), ("group_max", ctypes.c_uint32),
                ("wg_path_min", ctypes.c_uint64), ("clique_level_engine", ctypes.c_uint32),
                ("census_level_engine", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {"lane_path_max": int(self.lane_path_max), "item_grid": int(self.item_grid),
                "super_chunks": int(self.super_chunks), "batch_bytes": int(self.batch_bytes),
                "stream_ratio": int(self.stream_ratio), "group_max": int(self.group_max),
                "wg_path_min": int(self.wg_path_min),
                "clique_level_engine": int(self.clique_level_engine),
                "census_level_engine": int(self.census_level_engine)}


class _Trial(ctypes.Structure):
    _fields_ = [("tuning", _Tuning), ("ms", ctypes.c_double)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("mem_cap_bytes", ctypes.c_uint64), ("slots", ctypes.c_int32), ("merge", ctypes.c_int32),
                ("max_merge", ctypes.c_uint64), ("min_reserve_bytes", ctypes.c_uint64),
                ("hub_min_degree", ctypes.c_uint64), ("hub_max_bytes", ctypes.c_uint64),
                ("tuning", _Tuning)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("passes", ctypes.c_uint64), ("requests_done", ctypes.c_uint64),
                ("max_group", ctypes.c_uint64), ("peak_reserved", ctypes.c_uint64),
                ("peak_concurrent", ctypes.c_uint64), ("failed", ctypes.c_uint64)]


_lib = None


def library() -> ctypes.CDLL:
    """Loads libsme.so (built in-tree by __graft_entry__.build / make)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `make -C subgraphmatchgpu_amd`")
    lib = ctypes.CDLL(LIB_PATH)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    vp = ctypes.c_void_p
    sig = {
        "sme_open": ([ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)], ctypes.c_int),
        "sme_close": ([vp], None),
        "sme_last_error": ([vp], ctypes.c_char_p),
        "sme_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "sme_graph_upload": ([vp, vp, vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(vp)],
                             ctypes.c_int),
        "sme_graph_destroy": ([vp], None),
        "sme_plan_create": ([ctypes.c_int] * 4 + [ctypes.POINTER(vp)], ctypes.c_int),
        "sme_plan_describe": ([vp, ctypes.c_char_p, ctypes.c_size_t], ctypes.c_int),
        "sme_plan_levels": ([vp], ctypes.c_int),
        "sme_plan_set_reuse": ([vp, ctypes.c_int], ctypes.c_int),
        "sme_plan_destroy": ([vp], None),
        "sme_count": ([vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32, u64p, u64p,
                       ctypes.POINTER(_Stats)], ctypes.c_int),
        "sme_profile_enable": ([vp, ctypes.c_int], ctypes.c_int),
        "sme_profile_reset": ([vp], ctypes.c_int),
        "sme_profile_slots": ([], ctypes.c_int),
        "sme_profile_get": ([vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, u64p,
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), u64p],
                            ctypes.c_int),
        "sme_bin_read": ([ctypes.c_char_p, u64p, u64p, ctypes.POINTER(u64p), ctypes.POINTER(u32p)],
                         ctypes.c_int),
        "sme_bin_write": ([ctypes.c_char_p, vp, vp, ctypes.c_uint64, ctypes.c_uint64], ctypes.c_int),
        "sme_snap_read": ([ctypes.c_char_p, ctypes.c_int, u64p, u64p, ctypes.POINTER(u64p),
                           ctypes.POINTER(u32p)], ctypes.c_int),
        "sme_gen_rmat": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                          ctypes.c_double, ctypes.c_uint64, u64p, u64p, ctypes.POINTER(u64p),
                          ctypes.POINTER(u32p)], ctypes.c_int),
        "sme_gen_er": ([ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, u64p, u64p,
                        ctypes.POINTER(u64p), ctypes.POINTER(u32p)], ctypes.c_int),
        "sme_gen_er_batch": ([ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                              ctypes.c_double, ctypes.c_uint64, u64p, u64p, ctypes.POINTER(u64p),
                              ctypes.POINTER(u32p), ctypes.POINTER(u64p)], ctypes.c_int),
        "sme_free_host": ([vp], None),
        "sme_calibrate_copy": ([vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "sme_set_hub_bitmaps": ([vp, ctypes.c_uint64, ctypes.c_uint64], ctypes.c_int),
        "sme_count_buckets": ([vp, vp, vp, vp, ctypes.c_uint64, vp, u64p, u64p, ctypes.POINTER(_Stats)],
                              ctypes.c_int),
        "sme_graph_hubs": ([vp, u64p], ctypes.c_int),
        "sme_set_rank_order": ([vp, ctypes.c_int], ctypes.c_int),
        "sme_count_static": ([vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32, u64p, u64p,
                              ctypes.POINTER(_Stats)], ctypes.c_int),
        "sme_static_scratch_bytes": ([], ctypes.c_uint64),
        "sme_count_ranked": ([vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32, u64p, u64p,
                              ctypes.POINTER(_Stats)], ctypes.c_int),
        "sme_graph_ranked": ([vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "sme_set_tuning": ([vp, ctypes.POINTER(_Tuning)], ctypes.c_int),
        "sme_list_matches": ([vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.c_uint64, u64p, u64p],
                             ctypes.c_int),
        "sme_plan_vertices": ([vp], ctypes.c_int),
        "sme_server_open": ([ctypes.c_int, vp, vp, ctypes.c_uint64, ctypes.c_uint64,
                             ctypes.POINTER(_ServerConfig), ctypes.POINTER(vp)], ctypes.c_int),
        "sme_server_submit": ([vp, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, u64p], ctypes.c_int),
        "sme_server_wait": ([vp, ctypes.c_uint64, u64p, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "sme_server_get_stats": ([vp, ctypes.POINTER(_ServerStats)], ctypes.c_int),
        "sme_server_close": ([vp], None),
        "sme_get_tuning": ([vp, ctypes.POINTER(_Tuning)], ctypes.c_int),
        "sme_autotune": ([vp, vp, ctypes.POINTER(vp), ctypes.c_int, vp, ctypes.c_uint64,
                          ctypes.c_double, ctypes.POINTER(_Tuning), ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(_Trial), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
                         ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(rc: int, ctx=None):
    if rc != 0:
        msg = library().sme_last_error(ctx).decode(errors="replace")
        raise SmeError(rc, msg)


@dataclass
class Graph:
    """Host CSR in the reference's BIN layout (src/graph/GraphIO.h:104-127)."""
    row_ptrs: np.ndarray  # uint64[n+1]
    cols: np.ndarray      # uint32[nnz]

    @property
    def n(self) -> int:
        return int(self.row_ptrs.shape[0] - 1)

    @property
    def nnz(self) -> int:
        return int(self.cols.shape[0])

    def checksum(self) -> str:
        import hashlib
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(self.row_ptrs, dtype=np.uint64).tobytes())
        h.update(np.ascontiguousarray(self.cols, dtype=np.uint32).tobytes())
        return h.hexdigest()


def _take_host(n, nnz, rp, cols) -> Graph:
    lib = library()
    a = np.ctypeslib.as_array(rp, shape=(n.value + 1,)).copy()
    b = (np.ctypeslib.as_array(cols, shape=(nnz.value,)).copy() if nnz.value
         else np.zeros(0, dtype=np.uint32))
    lib.sme_free_host(ctypes.cast(rp, ctypes.c_void_p))
    lib.sme_free_host(ctypes.cast(cols, ctypes.c_void_p))
    return Graph(a.astype(np.uint64, copy=False), b.astype(np.uint32, copy=False))


def _out_args():
    return (ctypes.c_uint64(), ctypes.c_uint64(), ctypes.POINTER(ctypes.c_uint64)(),
            ctypes.POINTER(ctypes.c_uint32)())


def gen_rmat(scale: int, edge_factor: int, a=0.57, b=0.19, c=0.19, seed=1) -> Graph:
    """Seeded R-MAT, symmetrized, self loops and duplicates removed."""
    n, nnz, rp, cols = _out_args()
    _check(library().sme_gen_rmat(scale, edge_factor, a, b, c, seed, ctypes.byref(n),
                                  ctypes.byref(nnz), ctypes.byref(rp), ctypes.byref(cols)))
    return _take_host(n, nnz, rp, cols)


def gen_er(n_vertices: int, n_edges: int, seed=1) -> Graph:
    """Seeded G(n, m) with m distinct undirected edges."""
    n, nnz, rp, cols = _out_args()
    _check(library().sme_gen_er(n_vertices, n_edges, seed, ctypes.byref(n), ctypes.byref(nnz),
                                ctypes.byref(rp), ctypes.byref(cols)))
    return _take_host(n, nnz, rp, cols)


def gen_er_batch(ngraphs: int, n_range=(64, 8192), degree_range=(4.0, 48.0), seed=1):
    """`ngraphs` seeded G(n, m) graphs packed back to back (a disjoint union).

    Returns (graph, graph_offs): member k owns vertices
    [graph_offs[k], graph_offs[k+1]); n_k is uniform in n_range and the
    average degree uniform in degree_range."""
    n, nnz, rp, cols = _out_args()
    offs = ctypes.POINTER(ctypes.c_uint64)()
    _check(library().sme_gen_er_batch(ngraphs, n_range[0], n_range[1], float(degree_range[0]),
                                      float(degree_range[1]), seed, ctypes.byref(n),
                                      ctypes.byref(nnz), ctypes.byref(rp), ctypes.byref(cols),
                                      ctypes.byref(offs)))
    go = np.ctypeslib.as_array(offs, shape=(ngraphs + 1,)).copy()
    library().sme_free_host(ctypes.cast(offs, ctypes.c_void_p))
    return _take_host(n, nnz, rp, cols), go


def read_bin(path: str) -> Graph:
    n, nnz, rp, cols = _out_args()
    _check(library().sme_bin_read(path.encode(), ctypes.byref(n), ctypes.byref(nnz),
                                  ctypes.byref(rp), ctypes.byref(cols)))
    return _take_host(n, nnz, rp, cols)


def write_bin(path: str, g: Graph) -> None:
    rp = np.ascontiguousarray(g.row_ptrs, dtype=np.uint64)
    cols = np.ascontiguousarray(g.cols, dtype=np.uint32)
    _check(library().sme_bin_write(path.encode(), rp.ctypes.data, cols.ctypes.data, g.n, g.nnz))


def read_snap(path: str, directed: bool = False) -> Graph:
    n, nnz, rp, cols = _out_args()
    _check(library().sme_snap_read(path.encode(), int(directed), ctypes.byref(n), ctypes.byref(nnz),
                                   ctypes.byref(rp), ctypes.byref(cols)))
    return _take_host(n, nnz, rp, cols)


def from_adjacency(adj) -> Graph:
    """CSR from a list of neighbour lists (Graph(vector<vector<uintV>>),
    src/graph/CPUGraph.h:16-18); lists are sorted here."""
    rp = np.zeros(len(adj) + 1, dtype=np.uint64)
    for i, lst in enumerate(adj):
        rp[i + 1] = rp[i] + len(lst)
    cols = np.concatenate([np.sort(np.asarray(l, dtype=np.uint32)) for l in adj]) if len(adj) else \
        np.zeros(0, dtype=np.uint32)
    return Graph(rp, cols.astype(np.uint32))


class Plan:
    def __init__(self, query: int, algo: int = GPU_LIGHT, compress_level: int = CL_SPECIAL,
                 enable_ordering: bool = True, reuse: bool | None = None):
        lib = library()
        h = ctypes.c_void_p()
        _check(lib.sme_plan_create(algo, query, compress_level, int(enable_ordering), ctypes.byref(h)))
        self._h = h
        if reuse is not None:
            _check(lib.sme_plan_set_reuse(h, int(reuse)))
        self.query, self.algo, self.compress_level = query, algo, compress_level
        self.enable_ordering = enable_ordering

    @property
    def levels(self) -> int:
        return library().sme_plan_levels(self._h)

    def describe(self) -> str:
        lib = library()
        n = lib.sme_plan_describe(self._h, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib.sme_plan_describe(self._h, buf, n + 1)
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.sme_plan_destroy(self._h)
            self._h = None


@dataclass
class CountResult:
    count: int
    levels: list
    batches: int
    d2h_syncs: int
    peak_pool_bytes: int
    elapsed_ms: float
    ranked: bool = False


class Device:
    """One GPU: a stream plus a memory pool of `budget_bytes` for intermediates
    (the reference's per-device limit, src/Meta.h:26-28)."""

    def __init__(self, device: int = 0, budget_bytes: int = 8 << 30):
        lib = library()
        h = ctypes.c_void_p()
        _check(lib.sme_open(device, budget_bytes, ctypes.byref(h)))
        self._h = h
        self._graphs = []

    def upload(self, g: Graph) -> "DeviceGraph":
        return DeviceGraph(self, g)

    def count(self, dg: "DeviceGraph", plan: Plan, v_begin: int = 0, v_end: int | None = None) -> CountResult:
        lib = library()
        if v_end is None:
            v_end = dg.n
        cnt = ctypes.c_uint64()
        lv = (ctypes.c_uint64 * MAX_LEVELS)()
        st = _Stats()
        _check(lib.sme_count(self._h, dg._h, plan._h, v_begin, v_end, ctypes.byref(cnt), lv,
                             ctypes.byref(st)), self._h)
        return CountResult(int(cnt.value), [int(lv[i]) for i in range(plan.levels)], int(st.batches),
                           int(st.d2h_syncs), int(st.peak_pool_bytes), float(st.elapsed_ms),
                           bool(st.ranked))

    def count_static(self, dg: "DeviceGraph", plan: Plan, v_begin: int = 0, v_end: int | None = None) -> CountResult:
        """sme_count through the no-allocation entry point (intermediates in
        the device's static scratch array)."""
        lib = library()
        if v_end is None:
            v_end = dg.n
        cnt = ctypes.c_uint64()
        lv = (ctypes.c_uint64 * MAX_LEVELS)()
        st = _Stats()
        _check(lib.sme_count_static(self._h, dg._h, plan._h, v_begin, v_end, ctypes.byref(cnt), lv,
                                    ctypes.byref(st)), self._h)
        return CountResult(int(cnt.value), [int(lv[i]) for i in range(plan.levels)], int(st.batches),
                           int(st.d2h_syncs), int(st.peak_pool_bytes), float(st.elapsed_ms),
                           bool(st.ranked))

    def count_ranked(self, dg: "DeviceGraph", plan: Plan, r_begin: int, r_end: int,
                     levels: bool = True) -> CountResult:
        """Count over start range [r_begin, r_end) of the degree-ranked copy
        (ids of that copy); the tasks of a partition of [0, n) sum to the
        full count.  levels=False asks for no level outputs."""
        lib = library()
        cnt = ctypes.c_uint64()
        lv = (ctypes.c_uint64 * MAX_LEVELS)()
        st = _Stats()
        _check(lib.sme_count_ranked(self._h, dg._h, plan._h, r_begin, r_end, ctypes.byref(cnt),
                                    lv if levels else None, ctypes.byref(st)), self._h)
        return CountResult(int(cnt.value), [int(lv[i]) for i in range(plan.levels)] if levels else [],
                           int(st.batches), int(st.d2h_syncs), int(st.peak_pool_bytes),
                           float(st.elapsed_ms), bool(st.ranked))

    def count_only(self, dg: "DeviceGraph", plan: Plan, v_begin: int = 0, v_end: int | None = None) -> CountResult:
        """sme_count without level outputs (a full-range count may then run
        on the degree-ranked copy)."""
        lib = library()
        if v_end is None:
            v_end = dg.n
        cnt = ctypes.c_uint64()
        st = _Stats()
        _check(lib.sme_count(self._h, dg._h, plan._h, v_begin, v_end, ctypes.byref(cnt), None,
                             ctypes.byref(st)), self._h)
        return CountResult(int(cnt.value), [], int(st.batches), int(st.d2h_syncs),
                           int(st.peak_pool_bytes), float(st.elapsed_ms), bool(st.ranked))

    def count_buckets(self, dg: "DeviceGraph", plan: Plan, bounds) -> tuple:
        """One pass over [bounds[0], bounds[-1]); returns (per-bucket counts
        as uint64 array, CountResult of the whole pass)."""
        lib = library()
        b = np.ascontiguousarray(bounds, dtype=np.uint32)
        nb = b.size - 1
        out = np.zeros(max(nb, 1), dtype=np.uint64)
        cnt = ctypes.c_uint64()
        lv = (ctypes.c_uint64 * MAX_LEVELS)()
        st = _Stats()
        _check(lib.sme_count_buckets(self._h, dg._h, plan._h, b.ctypes.data, nb, out.ctypes.data,
                                     ctypes.byref(cnt), lv, ctypes.byref(st)), self._h)
        res = CountResult(int(cnt.value), [int(lv[i]) for i in range(plan.levels)], int(st.batches),
                          int(st.d2h_syncs), int(st.peak_pool_bytes), float(st.elapsed_ms))
        return out[:nb], res

    def profile(self, on: bool = True):
        _check(library().sme_profile_enable(self._h, int(on)), self._h)

    def profile_reset(self):
        _check(library().sme_profile_reset(self._h), self._h)

    def profile_read(self) -> dict:
        lib = library()
        out = {}
        for s in range(lib.sme_profile_slots()):
            name = ctypes.create_string_buffer(64)
            la, ms, by, it = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
            _check(lib.sme_profile_get(self._h, s, name, 64, ctypes.byref(la), ctypes.byref(ms),
                                       ctypes.byref(by), ctypes.byref(it)), self._h)
            out[name.value.decode()] = {"launches": la.value, "ms": ms.value, "bytes": by.value,
                                        "items": it.value}
        return out

    def set_hub_bitmaps(self, min_degree: int, max_bytes: int):
        """Adjacency bitmaps for vertices of degree >= min_degree in graphs
        uploaded afterwards (0 disables)."""
        _check(library().sme_set_hub_bitmaps(self._h, min_degree, max_bytes), self._h)

    def set_rank_order(self, on: bool = True):
        """Graphs uploaded afterwards keep a degree-ranked copy on which
        clique plans over all start vertices run (same count and levels)."""
        _check(library().sme_set_rank_order(self._h, int(on)), self._h)

    def list_matches(self, dg: "DeviceGraph", plan: Plan, v_begin: int = 0, v_end: int | None = None,
                     max_matches: int = 1 << 20):
        """Matches of a GPSM plan as an (rows, nv) uint32 array in
        lexicographic order (at most max_matches rows) and the full count."""
        lib = library()
        nv = lib.sme_plan_vertices(plan._h)
        out = np.zeros((max(max_matches, 1), nv), dtype=np.uint32)
        total, rows = ctypes.c_uint64(), ctypes.c_uint64()
        ve = dg.n if v_end is None else v_end
        _check(lib.sme_list_matches(self._h, dg._h, plan._h, v_begin, ve, out.ctypes.data, max_matches,
                                    ctypes.byref(total), ctypes.byref(rows)), self._h)
        return out[:rows.value], int(total.value)

    def set_tuning(self, lane_path_max: int = 0, item_grid: int = 0, batch_bytes: int = 0,
                   super_chunks: int = 0, stream_ratio: int = 0, group_max: int = 0,
                   wg_path_min: int = 0, clique_level_engine: int = 0, census_level_engine: int = 0):
        """Schedule tunables for later calls (results never depend on them):
        lane_path_max -- counting-level paths with at most this many pivot
        candidates are verified by one lane each; item_grid -- workgroup cap
        of the candidate kernels (0: default); batch_bytes -- cap on one
        batch's device bytes (0: the level's share of the budget);
        super_chunks -- 64-candidate chunks per wavefront step (runs inside
        one path load the path once and stage a list window in LDS);
        stream_ratio -- read an adjacency list in order as the pivot when it
        is at most this many times a short reused segment (0: off);
        group_max -- counting levels that extend a reused segment with one
        list count the paths sharing a segment of at most this many values
        in one wavefront, segment in LDS, lists read in order (0: off);
        wg_path_min -- counting-level paths with at least this many pivot
        candidates are counted by a whole workgroup each (0: off);
        clique_level_engine -- 1: clique plans run through the level engine
        instead of bitmap counting over each start vertex's upper
        neighbourhood (0, the default); census_level_engine -- 1: a house
        count over all start vertices without level outputs runs through the
        level engine instead of per-edge triangle counts and a 4-cycle wedge
        walk (0, the default)."""
        t = _Tuning(lane_path_max, item_grid, super_chunks, batch_bytes, stream_ratio, group_max,
                    wg_path_min, clique_level_engine, census_level_engine)
        _check(library().sme_set_tuning(self._h, ctypes.byref(t)), self._h)

    def get_tuning(self) -> dict:
        t = _Tuning()
        _check(library().sme_get_tuning(self._h, ctypes.byref(t)), self._h)
        return t.as_dict()

    def autotune(self, dg: "DeviceGraph", plans, bounds=None, sample_fraction: float = 0.02):
        """Times `plans` on a sample for each candidate setting, installs the
        fastest and returns (tuning, sample_ms, trials)."""
        plans = list(plans)
        arr = (ctypes.c_void_p * len(plans))(*[p._h for p in plans])
        if bounds is not None:
            b = np.ascontiguousarray(bounds, dtype=np.uint32)
            bp, nb = b.ctypes.data, b.size - 1
        else:
            b, bp, nb = None, None, 0
        best, ms, n = _Tuning(), ctypes.c_double(), ctypes.c_int()
        trials = (_Trial * 64)()
        _check(library().sme_autotune(self._h, dg._h, arr, len(plans), bp, nb, sample_fraction,
                                      ctypes.byref(best), ctypes.byref(ms), trials, 64,
                                      ctypes.byref(n)), self._h)
        log = [dict(trials[i].tuning.as_dict(), ms=float(trials[i].ms)) for i in range(min(n.value, 64))]
        return best.as_dict(), float(ms.value), log

    def calibrate_copy(self, nbytes: int) -> float:
        """One device-to-device copy of nbytes inside the pool; returns ms."""
        ms = ctypes.c_double()
        _check(library().sme_calibrate_copy(self._h, nbytes, ctypes.byref(ms)), self._h)
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            for g in self._graphs:
                g.release()
            library().sme_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraph:
    def __init__(self, dev: Device, g: Graph):
        lib = library()
        self._rp = np.ascontiguousarray(g.row_ptrs, dtype=np.uint64)
        self._cols = np.ascontiguousarray(g.cols, dtype=np.uint32)
        h = ctypes.c_void_p()
        _check(lib.sme_graph_upload(dev._h, self._rp.ctypes.data,
                                    self._cols.ctypes.data if g.nnz else None, g.n, g.nnz,
                                    ctypes.byref(h)), dev._h)
        self._h = h
        self.n, self.nnz = g.n, g.nnz
        dev._graphs.append(self)

    @property
    def hubs(self) -> int:
        c = ctypes.c_uint64()
        _check(library().sme_graph_hubs(self._h, ctypes.byref(c)))
        return c.value

    @property
    def ranked(self) -> bool:
        c = ctypes.c_int()
        _check(library().sme_graph_ranked(self._h, ctypes.byref(c)))
        return bool(c.value)

    def release(self):
        if getattr(self, "_h", None):
            library().sme_graph_destroy(self._h)
            self._h = None


def device_count() -> int:
    c = ctypes.c_int()
    _check(library().sme_device_count(ctypes.byref(c)))
    return c.value


class Server:
    """Pattern-count requests over one resident graph (see include/sme.h):
    continuous batching of same-pattern requests into bucketed passes,
    admission by estimated intermediate bytes under mem_cap_bytes, `slots`
    passes in flight.  merge=False, slots=1 is the plain one-at-a-time path."""

    def __init__(self, g: "Graph", mem_cap_bytes: int, slots: int = 4, merge: bool = True,
                 max_merge: int = 1024, min_reserve_bytes: int = 1 << 30, hub_min_degree: int = 0,
                 hub_max_bytes: int = 0, tuning: dict | None = None, device: int = 0):
        lib = library()
        t = tuning or {}
        cfg = _ServerConfig(mem_cap_bytes, slots, int(merge), max_merge, min_reserve_bytes,
                            hub_min_degree, hub_max_bytes,
                            _Tuning(t.get("lane_path_max", 0), t.get("item_grid", 0),
                                    t.get("super_chunks", 0), t.get("batch_bytes", 0),
                                    t.get("stream_ratio", 0), t.get("group_max", 0),
                                    t.get("wg_path_min", 0), t.get("clique_level_engine", 0),
                                    t.get("census_level_engine", 0)))
        self._rp = np.ascontiguousarray(g.row_ptrs, dtype=np.uint64)
        self._cols = np.ascontiguousarray(g.cols, dtype=np.uint32)
        h = ctypes.c_void_p()
        _check(lib.sme_server_open(device, self._rp.ctypes.data, self._cols.ctypes.data, g.n, g.nnz,
                                   ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.n = g.n

    def submit(self, query: int, v_begin: int = 0, v_end: int | None = None) -> int:
        rid = ctypes.c_uint64()
        _check(library().sme_server_submit(self._h, query, v_begin, self.n if v_end is None else v_end,
                                           ctypes.byref(rid)))
        return rid.value

    def wait(self, rid: int) -> tuple:
        """(count, latency_ms) of request rid, blocking until it is done."""
        c, ms = ctypes.c_uint64(), ctypes.c_double()
        _check(library().sme_server_wait(self._h, rid, ctypes.byref(c), ctypes.byref(ms)))
        return int(c.value), float(ms.value)

    def stats(self) -> dict:
        st = _ServerStats()
        _check(library().sme_server_get_stats(self._h, ctypes.byref(st)))
        return {k: int(getattr(st, k)) for k, _ in _ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            library().sme_server_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
