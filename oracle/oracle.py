"""ctypes view of liboracle.so -- the CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker.  See oracle.h for what
each function restates (reference file:line) and oracle.c for how each one is
pinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_CPU = os.path.join(REF_DIR, "patternmatchcpu")
REF_PLAN_DUMP = os.path.join(REF_DIR, "plan_dump")
MAX_OPS = 16

_lib = None


def build() -> None:
    subprocess.run(["make", "-C", HERE, "-s", "liboracle.so"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        u64 = ctypes.c_uint64
        L.or_light_plan.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        L.or_plan_text.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.or_wcoj_count.argtypes = [u64, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(u64)]
        L.or_light_count.argtypes = [u64, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                     ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.or_gpsm_count.argtypes = [u64, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(u64),
                                    ctypes.POINTER(u64)]
        _lib = L
    return _lib


def plan_text(q: int, compress_level: int = 3, enable_ordering: bool = True) -> str:
    L = lib()
    p = ctypes.create_string_buffer(16384)
    if L.or_light_plan(q, compress_level, int(enable_ordering), p) != 0:
        raise ValueError("bad query / compress level")
    buf = ctypes.create_string_buffer(16384)
    L.or_plan_text(p, buf, 16384)
    return buf.value.decode()


def _arrays(rp, cols):
    rp = np.ascontiguousarray(rp, dtype=np.uint64)
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    if cols.size == 0:
        cols = np.zeros(1, dtype=np.uint32)
    return rp, cols


def wcoj_count(rp, cols, q, enable_ordering=True, v_begin=0, v_end=None, threads=0) -> int:
    rp, cols = _arrays(rp, cols)
    n = rp.size - 1
    out = ctypes.c_uint64()
    rc = lib().or_wcoj_count(n, rp.ctypes.data, cols.ctypes.data, q, int(enable_ordering), v_begin,
                             n if v_end is None else v_end, threads, ctypes.byref(out))
    if rc:
        raise ValueError(rc)
    return int(out.value)


def wcoj_count_list(rp, cols, q, starts, enable_ordering=True, threads=0) -> int:
    """CPU_WCOJ count over an explicit list of start vertices."""
    rp, cols = _arrays(rp, cols)
    starts = np.ascontiguousarray(starts, dtype=np.uint32)
    L = lib()
    L.or_wcoj_count_list.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_uint64)]
    out = ctypes.c_uint64()
    rc = L.or_wcoj_count_list(rp.size - 1, rp.ctypes.data, cols.ctypes.data, q, int(enable_ordering),
                              starts.ctypes.data, starts.size, threads, ctypes.byref(out))
    if rc:
        raise ValueError(rc)
    return int(out.value)


def light_count(rp, cols, q, compress_level=3, enable_ordering=True, v_begin=0, v_end=None,
                threads=0):
    """Returns (count, per-exec-level outputs)."""
    rp, cols = _arrays(rp, cols)
    n = rp.size - 1
    out = ctypes.c_uint64()
    lv = (ctypes.c_uint64 * MAX_OPS)()
    rc = lib().or_light_count(n, rp.ctypes.data, cols.ctypes.data, q, compress_level,
                              int(enable_ordering), v_begin, n if v_end is None else v_end, threads,
                              ctypes.byref(out), lv)
    if rc:
        raise ValueError(rc)
    nops = plan_text(q, compress_level, enable_ordering).split("\n")[1].count(" ")
    return int(out.value), [int(lv[i]) for i in range(nops)]


def light_count_per_vertex(rp, cols, q, v_begin=0, v_end=None, compress_level=3, enable_ordering=True,
                           threads=0):
    """LIGHT count of every start vertex in [v_begin, v_end) (uint64 array)."""
    rp, cols = _arrays(rp, cols)
    n = rp.size - 1
    v_end = n if v_end is None else v_end
    out = np.zeros(max(0, v_end - v_begin), dtype=np.uint64)
    L = lib()
    L.or_light_count_per_vertex.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_int, ctypes.c_void_p]
    if out.size:
        rc = L.or_light_count_per_vertex(n, rp.ctypes.data, cols.ctypes.data, q, compress_level,
                                         int(enable_ordering), v_begin, v_end, threads, out.ctypes.data)
        if rc:
            raise ValueError(rc)
    return out


def gpsm_count(rp, cols, q, enable_ordering=True, v_begin=0, v_end=None, threads=0):
    rp, cols = _arrays(rp, cols)
    n = rp.size - 1
    out = ctypes.c_uint64()
    lv = (ctypes.c_uint64 * MAX_OPS)()
    rc = lib().or_gpsm_count(n, rp.ctypes.data, cols.ctypes.data, q, int(enable_ordering), v_begin,
                             n if v_end is None else v_end, threads, ctypes.byref(out), lv)
    if rc:
        raise ValueError(rc)
    return int(out.value), [int(lv[i]) for i in range(MAX_OPS)]


class _Census(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("edges", "self_loops", "triangles", "diamonds", "k4", "k5",
                                               "max_out_degree")]


def census(rp, cols, kmax=4, threads=0) -> dict:
    """Whole-graph counts by degree-ordered clique listing (census.c), an
    independent check of the plan-following functions: edges, triangles
    (Q0), diamonds (Q2), 4-cliques (Q3, kmax >= 4), 5-cliques (Q7, kmax 5)."""
    rp, cols = _arrays(rp, cols)
    L = lib()
    L.or_census.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                            ctypes.POINTER(_Census)]
    out = _Census()
    rc = L.or_census(rp.size - 1, rp.ctypes.data, cols.ctypes.data, kmax, threads, ctypes.byref(out))
    if rc:
        raise ValueError(rc)
    return {k: int(getattr(out, k)) for k, _ in _Census._fields_}


def ref_available() -> bool:
    return os.path.exists(REF_CPU)


def ref_count(bin_path: str, q: int, algo: int = 0, threads: int = 1, buffer_gb: int = 1,
              timeout: float = 600.0):
    """Runs the reference CPU binary (built from the reference sources into
    oracle/_ref) and returns (total_match_count, elapsed_ms)."""
    cmd = [REF_CPU, "-f", bin_path, "-d", "0", "-a", str(algo), "-q", str(q), "-t", str(threads),
           "-b", str(buffer_gb)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    for line in r.stdout.splitlines():
        if line.startswith("total_match_count="):
            a, b = line.split(",")
            return int(a.split("=")[1]), float(b.split("=")[1].rstrip("ms"))
    raise RuntimeError(f"reference run failed: rc={r.returncode} {r.stdout[-400:]} {r.stderr[-400:]}")


def gpsm_list(rp, cols, q, v_begin=0, v_end=None, cap=1 << 20, enable_ordering=True):
    """GPSM matches as an (rows, nv) uint32 array (lexicographic) and the
    full count."""
    rp, cols = _arrays(rp, cols)
    n = rp.size - 1
    nv = (3, 4, 4, 4, 5, 6, 5, 5, 5, 5, 5, 6, 5, 6)[q]
    out = np.zeros((max(cap, 1), nv), dtype=np.uint32)
    rows, total = ctypes.c_uint64(), ctypes.c_uint64()
    L = lib()
    L.or_gpsm_list.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                               ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64,
                               ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    rc = L.or_gpsm_list(n, rp.ctypes.data, cols.ctypes.data, q, int(enable_ordering), v_begin,
                        n if v_end is None else v_end, out.ctypes.data, cap, ctypes.byref(rows),
                        ctypes.byref(total))
    if rc:
        raise ValueError(rc)
    return out[:rows.value], int(total.value)
