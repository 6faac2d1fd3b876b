"""ctypes wrapper of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker and the CPU timing baseline -- never by the product package.

* ``liboracle.so``   the C restatement in this folder (always built).
* ``_ref/libref.so`` the reference's own csr.cpp + cpuBFS.cpp compiled by
  ``make ref`` (only where the reference sources are present; the built file
  travels with the tree).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

_lib = None
_ref = None

P = C.c_void_p
i64p = C.POINTER(C.c_int64)


def build(ref=True):
    subprocess.run(["make", "-C", HERE, "-B", "liboracle.so"], check=True, capture_output=True)
    if ref:
        subprocess.run(["make", "-C", HERE, "-B", "ref"], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        h = C.CDLL(LIB)
        h.oracle_load_edgelist.argtypes = [C.c_char_p, C.c_int, i64p, i64p, P, P]
        h.oracle_bfs.argtypes = [C.c_int64, P, P, C.c_int32, P, P, i64p, i64p, C.POINTER(C.c_int32)]
        h.oracle_bfs_timed.argtypes = [C.c_int64, P, P, C.c_int32, P, C.POINTER(C.c_double), i64p, i64p]
        h.oracle_bfs_dequeue.argtypes = [C.c_int64, P, P, C.c_int32, P]
        h.oracle_bfs_multi.argtypes = [C.c_int64, P, P, P, C.c_int, P]
        h.oracle_check_levels.argtypes = [C.c_int64, P, P, C.c_int32, P, i64p]
        h.oracle_check_parents.argtypes = [C.c_int64, P, P, C.c_int32, P, P, i64p]
        h.oracle_exclusive_scan_u32.argtypes = [P, P, C.c_int64]
        h.oracle_exclusive_scan_u32.restype = None
        h.oracle_blelloch_scan_u32.argtypes = [P, C.c_int64]
        h.oracle_serialized_bfs.argtypes = [C.c_int64, P, P, C.c_int32, P, P, C.POINTER(C.c_int32)]
        _lib = h
    return _lib


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(REF_LIB)
        h = C.CDLL(REF_LIB)
        ip = C.POINTER(C.c_int)
        h.ref_load_csr.argtypes = [C.c_char_p, ip, ip, P, P]
        h.ref_cpu_bfs.argtypes = [C.c_int, P, P, C.c_int, P, C.POINTER(C.c_float), C.POINTER(C.c_double), ip]
        _ref = h
    return _ref


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _csr(offsets, cols):
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    if cols.size == 0:
        cols = np.zeros(1, dtype=np.int32)
    return offsets, cols


class OracleError(RuntimeError):
    pass


def load_edgelist(path, one_based=False):
    n, m = C.c_int64(), C.c_int64()
    rc = lib().oracle_load_edgelist(str(path).encode(), int(one_based), C.byref(n), C.byref(m), None, None)
    if rc:
        raise OracleError(f"oracle_load_edgelist: {rc}")
    off = np.zeros(n.value + 1, dtype=np.int64)
    cols = np.zeros(max(m.value, 1), dtype=np.int32)
    rc = lib().oracle_load_edgelist(str(path).encode(), int(one_based), C.byref(n), C.byref(m), _p(off), _p(cols))
    if rc:
        raise OracleError(f"oracle_load_edgelist: {rc}")
    return off, cols[: m.value]


def bfs(offsets, cols, source, parents=False):
    """Returns (levels, parents or None, stats)."""
    offsets, cols = _csr(offsets, cols)
    n = offsets.size - 1
    levels = np.empty(n, dtype=np.int32)
    par = np.empty(n, dtype=np.int32) if parents else None
    nr, mr, nl = C.c_int64(), C.c_int64(), C.c_int32()
    rc = lib().oracle_bfs(n, _p(offsets), _p(cols), int(source), _p(levels), _p(par), C.byref(nr), C.byref(mr),
                          C.byref(nl))
    if rc:
        raise OracleError(f"oracle_bfs: {rc}")
    return levels, par, {"n_reached": nr.value, "m_reached": mr.value, "num_levels": nl.value}


def bfs_timed(offsets, cols, source):
    """Returns (levels, ms, n_reached, m_reached) with the reference's timing region."""
    offsets, cols = _csr(offsets, cols)
    n = offsets.size - 1
    levels = np.empty(n, dtype=np.int32)
    ms, nr, mr = C.c_double(), C.c_int64(), C.c_int64()
    rc = lib().oracle_bfs_timed(n, _p(offsets), _p(cols), int(source), _p(levels), C.byref(ms), C.byref(nr),
                                C.byref(mr))
    if rc:
        raise OracleError(f"oracle_bfs_timed: {rc}")
    return levels, ms.value, nr.value, mr.value


def bfs_dequeue(offsets, cols, source):
    offsets, cols = _csr(offsets, cols)
    levels = np.empty(offsets.size - 1, dtype=np.int32)
    rc = lib().oracle_bfs_dequeue(offsets.size - 1, _p(offsets), _p(cols), int(source), _p(levels))
    if rc:
        raise OracleError(f"oracle_bfs_dequeue: {rc}")
    return levels


def bfs_multi(offsets, cols, sources):
    offsets, cols = _csr(offsets, cols)
    sources = np.ascontiguousarray(sources, dtype=np.int32)
    n = offsets.size - 1
    levels = np.empty((sources.size, n), dtype=np.int32)
    rc = lib().oracle_bfs_multi(n, _p(offsets), _p(cols), _p(sources), sources.size, _p(levels))
    if rc:
        raise OracleError(f"oracle_bfs_multi: {rc}")
    return levels


def check_levels(offsets, cols, source, levels):
    """0 if `levels` is the BFS distance from `source`, else (code, vertex)."""
    offsets, cols = _csr(offsets, cols)
    levels = np.ascontiguousarray(levels, dtype=np.int32)
    where = C.c_int64(-1)
    rc = lib().oracle_check_levels(offsets.size - 1, _p(offsets), _p(cols), int(source), _p(levels),
                                   C.byref(where))
    return 0 if rc == 0 else (rc, where.value)


def check_parents(offsets, cols, source, levels, parents):
    offsets, cols = _csr(offsets, cols)
    levels = np.ascontiguousarray(levels, dtype=np.int32)
    parents = np.ascontiguousarray(parents, dtype=np.int32)
    where = C.c_int64(-1)
    rc = lib().oracle_check_parents(offsets.size - 1, _p(offsets), _p(cols), int(source), _p(levels),
                                    _p(parents), C.byref(where))
    return 0 if rc == 0 else (rc, where.value)


def exclusive_scan(values):
    v = np.ascontiguousarray(values, dtype=np.int32).view(np.uint32)
    out = np.empty_like(v)
    lib().oracle_exclusive_scan_u32(_p(v), _p(out), v.size)
    return out.view(np.int32)


def blelloch_scan(values):
    v = np.array(values, dtype=np.uint32)
    rc = lib().oracle_blelloch_scan_u32(_p(v), v.size)
    if rc:
        raise OracleError("length must be a power of two")
    return v.view(np.int32)


def serialized_bfs(offsets, cols, source):
    """Returns (levels, frontier sizes per level)."""
    offsets, cols = _csr(offsets, cols)
    n = offsets.size - 1
    levels = np.empty(n, dtype=np.int32)
    fs = np.zeros(n + 2, dtype=np.int64)
    nl = C.c_int32()
    rc = lib().oracle_serialized_bfs(n, _p(offsets), _p(cols), int(source), _p(levels), _p(fs), C.byref(nl))
    if rc:
        raise OracleError(f"oracle_serialized_bfs: {rc}")
    return levels, fs[: nl.value]


# -- the reference's own code (oracle/_ref/libref.so) -----------------------

def ref_load_csr(path):
    """csr::csr on `path`: returns (row_offset int32[n+1], col_idx int32[m])."""
    n, m = C.c_int(), C.c_int()
    rc = ref().ref_load_csr(str(path).encode(), C.byref(n), C.byref(m), None, None)
    if rc:
        raise OracleError(f"ref_load_csr: {rc}")
    ro = np.zeros(n.value + 1, dtype=np.int32)
    ci = np.zeros(max(m.value, 1), dtype=np.int32)
    rc = ref().ref_load_csr(str(path).encode(), C.byref(n), C.byref(m), _p(ro), _p(ci))
    if rc:
        raise OracleError(f"ref_load_csr: {rc}")
    return ro, ci[: m.value]


def ref_cpu_bfs(offsets, cols, source):
    """cpuBFS::cpuBFS: returns (levels int32 with -1, clock() ms, wall ms, iterations)."""
    ro = np.ascontiguousarray(offsets, dtype=np.int64)
    if ro[-1] >= 2**31:
        raise OracleError("the reference CSR uses int offsets")
    ro = ro.astype(np.int32)
    ci = np.ascontiguousarray(cols, dtype=np.int32)
    if ci.size == 0:
        ci = np.zeros(1, dtype=np.int32)
    n = ro.size - 1
    levels = np.empty(n, dtype=np.int32)
    ms, wall, it = C.c_float(), C.c_double(), C.c_int()
    rc = ref().ref_cpu_bfs(n, _p(ro), _p(ci), int(source), _p(levels), C.byref(ms), C.byref(wall), C.byref(it))
    if rc:
        raise OracleError(f"ref_cpu_bfs: {rc}")
    return levels, ms.value, wall.value, it.value
