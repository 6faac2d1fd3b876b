"""MI355X (gfx950) BFS engine: Python side of the C ABI.

Mirrors the reference's host interface (YKMorsy/GPU_Graphs, `1D Partition`):

* ``csr(filename)``       -- csr::csr (include/csr.h:9-23, src/csr.cpp:3-67)
  with ``row_offset``, ``col_idx``, ``num_nodes``, ``num_edges``,
  ``print_info()``.
* ``gpuBFS(graph, source)`` -- gpuBFS::gpuBFS (include/gpuBFS.cuh:9-20,
  src/gpuBFS.cu:13-116): runs the whole BFS in the constructor and exposes
  ``host_distance``, ``exec_time`` (ms), ``iteration``,
  ``total_edges_traversed`` and ``print_distance()``.

Differences that the reference leaves undefined or gets wrong are fixed here:
unreached vertices are -1 (not INT_MAX), ids are bounds-checked, and the
statistics are real (the reference's are always 0, gpuBFS_kernels.cu:37,
gpuBFS.cu:63).  Everything runs through libgpugraphs.so; there is no CPU
path.
"""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import (GB_AUTO_N, GB_DROP_SELF_LOOPS, GB_EAGAIN, GB_EBUSY, GB_EDEVICE, GB_EINVAL, GB_EIO, GB_ENOMEM,
                   GB_ERANGE, GB_KEEP_ORDER, GB_MODE_ASYNC, GB_MODE_GRAPH, GB_MODE_LEVEL_LOOP, GB_OK, GB_ONE_BASED, GB_SYMMETRIZE,
                   GB_DEEP_AUTO, GB_DEEP_OFF, GB_DEEP_ON, GB_TAIL_AUTO, GB_TAIL_DECLINED, GB_TAIL_FINISHED,
                   GB_TAIL_HANDED_BACK, GB_TAIL_OFF, BatchStats, GBError, Stats, check, lib)

GB_DEGREE_BUCKETS = 34

__all__ = [
    "Graph", "DeviceGraph", "DeviceArray", "Server", "csr", "gpuBFS", "GBError", "Stats", "device_count",
    "exclusive_scan", "GB_AUTO_N", "GB_ONE_BASED", "GB_SYMMETRIZE", "GB_KEEP_ORDER", "GB_DROP_SELF_LOOPS",
    "GB_MODE_ASYNC", "GB_MODE_LEVEL_LOOP", "GB_MODE_GRAPH", "GB_OK", "GB_EINVAL", "GB_EIO", "GB_ENOMEM", "GB_EDEVICE", "GB_ERANGE",
    "GB_EBUSY", "GB_EAGAIN", "GB_DEEP_AUTO", "GB_DEEP_ON", "GB_DEEP_OFF", "GB_TAIL_AUTO", "GB_TAIL_OFF",
    "GB_TAIL_FINISHED", "GB_TAIL_DECLINED", "GB_TAIL_HANDED_BACK",
]


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def device_count():
    c = C.c_int(0)
    check(lib().gb_device_count(C.byref(c)))
    return c.value


class Graph:
    """Host CSR (int64 offsets, int32 columns) owned by the library."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    # -- constructors ----------------------------------------------------
    @classmethod
    def from_file(cls, path, flags=0):
        h = C.c_void_p()
        check(lib().gb_graph_load_edgelist(str(path).encode(), flags, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, n, src, dst, flags=0):
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        if src.shape != dst.shape:
            raise ValueError("src and dst differ in length")
        h = C.c_void_p()
        check(lib().gb_graph_from_edges(int(n), int(src.size), _ptr(src), _ptr(dst), flags, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csr(cls, offsets, cols, flags=0):
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        h = C.c_void_p()
        check(lib().gb_graph_from_csr(int(offsets.size - 1), _ptr(offsets), _ptr(cols), flags, C.byref(h)))
        return cls(h.value)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, permute=True):
        h = C.c_void_p()
        check(lib().gb_graph_generate_rmat(scale, edge_factor, a, b, c, seed, int(permute), C.byref(h)))
        return cls(h.value)

    @classmethod
    def erdos_renyi(cls, n_candidates, n_isolated, m_undirected, seed=1):
        h = C.c_void_p()
        check(lib().gb_graph_generate_er(n_candidates, n_isolated, m_undirected, seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def grid(cls, width, height, extra_fraction=0.05, seed=1):
        h = C.c_void_p()
        check(lib().gb_graph_generate_grid(width, height, extra_fraction, seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def load_binary(cls, path):
        h = C.c_void_p()
        check(lib().gb_graph_load_binary(str(path).encode(), C.byref(h)))
        return cls(h.value)

    # -- accessors -------------------------------------------------------
    @property
    def n(self):
        n = C.c_int64()
        check(lib().gb_graph_info(self._h, C.byref(n), None))
        return n.value

    @property
    def m(self):
        m = C.c_int64()
        check(lib().gb_graph_info(self._h, None, C.byref(m)))
        return m.value

    @property
    def offsets(self):
        """Borrowed int64[n+1] view (valid while this Graph lives)."""
        o, c = C.c_void_p(), C.c_void_p()
        check(lib().gb_graph_csr(self._h, C.byref(o), C.byref(c)))
        arr = np.ctypeslib.as_array(C.cast(o, C.POINTER(C.c_int64)), shape=(self.n + 1,))
        arr.flags.writeable = False
        return arr

    @property
    def cols(self):
        """Borrowed int32[m] view (valid while this Graph lives)."""
        o, c = C.c_void_p(), C.c_void_p()
        check(lib().gb_graph_csr(self._h, C.byref(o), C.byref(c)))
        m = self.m
        if m == 0:
            return np.zeros(0, dtype=np.int32)
        arr = np.ctypeslib.as_array(C.cast(c, C.POINTER(C.c_int32)), shape=(m,))
        arr.flags.writeable = False
        return arr

    def degrees(self):
        return np.diff(self.offsets)

    def degree_buckets(self):
        """(offsets[35], vertices[n]): bucket 0 = degree 0, bucket b >= 1 =
        degree in [2^(b-1), 2^b), ids ascending inside a bucket."""
        off = np.zeros(GB_DEGREE_BUCKETS + 1, dtype=np.int64)
        v = C.c_void_p()
        check(lib().gb_graph_degree_buckets(self._h, off.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(v)))
        n = self.n
        if n == 0:
            return off, np.zeros(0, dtype=np.int32)
        arr = np.ctypeslib.as_array(C.cast(v, C.POINTER(C.c_int32)), shape=(n,)).copy()
        return off, arr

    @property
    def is_symmetric(self):
        s = C.c_int()
        check(lib().gb_graph_is_symmetric(self._h, C.byref(s)))
        return bool(s.value)

    def write_edgelist(self, path):
        check(lib().gb_graph_write_edgelist(self._h, str(path).encode()))

    def save_binary(self, path):
        check(lib().gb_graph_save_binary(self._h, str(path).encode()))

    def free(self):
        if self._h is not None and self._h.value:
            lib().gb_graph_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """Plain device allocation (gb_dmalloc) of a numpy dtype and length."""

    def __init__(self, device, dtype, count):
        self.device = device
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        p = C.c_void_p()
        check(lib().gb_dmalloc(device, self.dtype.itemsize * max(self.count, 1), C.byref(p)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, device, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(device, arr.dtype, arr.size)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        if arr.size != self.count:
            raise ValueError("size mismatch")
        check(lib().gb_memcpy_htod(self.device, self.ptr, _ptr(arr), arr.nbytes))

    def numpy(self):
        out = np.empty(self.count, dtype=self.dtype)
        check(lib().gb_memcpy_dtoh(self.device, _ptr(out), self.ptr, out.nbytes))
        return out

    def numpy_slice(self, start, count):
        """Copies elements [start, start + count) to a new host array."""
        start, count = int(start), int(count)
        if start < 0 or count < 0 or start + count > self.count:
            raise ValueError("slice outside the array")
        out = np.empty(count, dtype=self.dtype)
        src = C.c_void_p(self.ptr.value + start * self.dtype.itemsize)
        check(lib().gb_memcpy_dtoh(self.device, _ptr(out), src, out.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().gb_dfree(self.device, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceGraph:
    """A graph resident in the HBM of one GPU."""

    def __init__(self, graph, device=0):
        h = C.c_void_p()
        check(lib().gb_device_upload(graph._h, device, C.byref(h)))
        self._h = h
        self.device = device
        self.n = graph.n
        self.m = graph.m

    def config(self, block=256, edges_per_thread=4, blocks_per_cu=0, mode=GB_MODE_ASYNC, dense_edges=0):
        """Kernel tunables; dense_edges=0 selects the default threshold."""
        check(lib().gb_bfs_config(self._h, block, edges_per_thread, blocks_per_cu, mode, dense_edges))

    def get_config(self):
        b, k, bpc, mode = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        dense = C.c_int64()
        check(lib().gb_bfs_get_config(self._h, C.byref(b), C.byref(k), C.byref(bpc), C.byref(mode), C.byref(dense)))
        return {"block": b.value, "edges_per_thread": k.value, "blocks_per_cu": bpc.value, "mode": mode.value,
                "dense_edges": dense.value}

    def direction_config(self, enable=True, alpha=0, beta=0):
        """Bottom-up levels (symmetric graphs only); alpha/beta 0 keep the current values."""
        check(lib().gb_bfs_direction_config(self._h, int(enable), int(alpha), int(beta)))

    def get_direction_config(self):
        e, a, b, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().gb_bfs_get_direction_config(self._h, C.byref(e), C.byref(a), C.byref(b), C.byref(s)))
        return {"enable": bool(e.value), "alpha": a.value, "beta": b.value, "symmetric": bool(s.value)}

    def deep_config(self, mode=GB_DEEP_AUTO):
        """One launch per level for deep searches with thin frontiers (GB_DEEP_*)."""
        check(lib().gb_bfs_deep_config(self._h, int(mode)))

    def get_deep_config(self):
        """{"mode": GB_DEEP_*, "used": whether the last bfs ran in deep mode}."""
        m, u = C.c_int(), C.c_int()
        check(lib().gb_bfs_get_deep_config(self._h, C.byref(m), C.byref(u)))
        return {"mode": m.value, "used": bool(u.value)}

    def tail_config(self, level=GB_TAIL_AUTO):
        """Level from which one workgroup runs the rest of a search: GB_TAIL_AUTO
        (learned), GB_TAIL_OFF, or a level >= 0."""
        check(lib().gb_bfs_tail_config(self._h, int(level)))

    def get_tail_config(self):
        """{"level": the setting, "used": level the last bfs started its tail
        at (-1: none), "outcome": GB_TAIL_FINISHED / DECLINED / HANDED_BACK, 0 if none}."""
        lv, u, o = C.c_int(), C.c_int(), C.c_int()
        check(lib().gb_bfs_get_tail_config(self._h, C.byref(lv), C.byref(u), C.byref(o)))
        return {"level": lv.value, "used": u.value, "outcome": o.value}

    def launch_hints(self):
        """(depth, bu_lo, bu_hi): levels queued before the first done check and
        the levels that get a bottom-up launch, learned from earlier searches."""
        d, lo, hi = C.c_int(), C.c_int(), C.c_int()
        check(lib().gb_bfs_launch_hints(self._h, C.byref(d), C.byref(lo), C.byref(hi)))
        return d.value, lo.value, hi.value

    def last_bottom_up(self):
        """(first, last) level the last bfs() searched bottom-up, (-1, -1) if none."""
        f, l = C.c_int(), C.c_int()
        check(lib().gb_bfs_last_bottom_up(self._h, C.byref(f), C.byref(l)))
        return f.value, l.value

    def memory_bytes(self):
        b = C.c_int64()
        check(lib().gb_dgraph_memory(self._h, C.byref(b)))
        return b.value

    def bfs(self, source, parents=False):
        """Host-pointer BFS: returns (levels, parents or None, stats dict)."""
        levels = np.empty(self.n, dtype=np.int32)
        par = np.empty(self.n, dtype=np.int32) if parents else None
        st = Stats()
        check(lib().gb_bfs(self._h, int(source), _ptr(levels), _ptr(par), C.byref(st)))
        return levels, par, st.as_dict()

    def bfs_device(self, source, d_levels, d_parents=None, raw=False):
        """Device-pointer BFS: results stay in HBM (DeviceArray outputs).
        Returns the statistics as a dict, or (raw=True) as the gb_stats
        structure itself (no conversion inside a timed loop)."""
        st = Stats()
        check(lib().gb_bfs_device(self._h, int(source), d_levels.ptr, d_parents.ptr if d_parents else None,
                                  C.byref(st)))
        return st if raw else st.as_dict()

    def batch_config(self, block=256, edges_per_thread=2, use_graph=True, sparse_edges=0):
        check(lib().gb_batch_config(self._h, block, edges_per_thread, int(use_graph), int(sparse_edges)))

    def get_batch_config(self):
        b, k, g, s = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        check(lib().gb_batch_get_config(self._h, C.byref(b), C.byref(k), C.byref(g), C.byref(s)))
        return {"block": b.value, "edges_per_thread": k.value, "use_graph": bool(g.value),
                "sparse_edges": s.value}

    def batch_direction_config(self, enable=True, alpha=0, beta=-1):
        """Bottom-up levels of bfs_batch; alpha 0 and beta -1 keep the current values."""
        check(lib().gb_batch_direction_config(self._h, int(enable), int(alpha), int(beta)))

    def get_batch_direction_config(self):
        e, a, b = C.c_int(), C.c_int(), C.c_int()
        check(lib().gb_batch_get_direction_config(self._h, C.byref(e), C.byref(a), C.byref(b)))
        return {"enable": bool(e.value), "alpha": a.value, "beta": b.value}

    def batch_deep_config(self, mode=GB_DEEP_AUTO):
        """Fused launches per level for deep batches with thin frontiers (GB_DEEP_*)."""
        check(lib().gb_batch_deep_config(self._h, int(mode)))

    def get_batch_deep_config(self):
        """{"mode": GB_DEEP_*, "used": whether the last batch ran fused levels}."""
        m, u = C.c_int(), C.c_int()
        check(lib().gb_batch_get_deep_config(self._h, C.byref(m), C.byref(u)))
        return {"mode": m.value, "used": bool(u.value)}

    def batch_last_bottom_up(self):
        """Number of levels the last bfs_batch ran bottom-up."""
        v = C.c_int()
        check(lib().gb_batch_last_bottom_up(self._h, C.byref(v)))
        return v.value

    def autotune(self, sources, reps=3):
        """Tunes bfs() over the given sources (1..64) on the device; the
        fastest configuration stays on the handle.  Returns the result dict
        (default_ms, best_ms, tried and the chosen values)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        res = _ffi.TuneResult()
        check(lib().gb_bfs_autotune(self._h, _ptr(src), src.size, int(reps), C.byref(res)))
        return res.as_dict()

    def batch_autotune(self, sources, reps=3):
        """Tunes bfs_batch() on one batch of sources (1..64)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        res = _ffi.TuneResult()
        check(lib().gb_batch_autotune(self._h, _ptr(src), src.size, int(reps), C.byref(res)))
        return res.as_dict()

    @staticmethod
    def _tune_struct(t):
        res = _ffi.TuneResult()
        for k, _ in res._fields_:
            if k in t:
                setattr(res, k, t[k])
        return res

    def apply_tuning(self, t):
        """Applies a dict returned by autotune() (possibly from another process)."""
        check(lib().gb_bfs_apply_tuning(self._h, C.byref(self._tune_struct(t))))

    def apply_batch_tuning(self, t):
        """Applies a dict returned by batch_autotune()."""
        check(lib().gb_batch_apply_tuning(self._h, C.byref(self._tune_struct(t))))

    def bfs_batch(self, sources, stats=True):
        """Multi-source BFS of up to 64 sources: returns (levels[k, n], stats or None)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        levels = np.empty((src.size, self.n), dtype=np.int32)
        st = BatchStats() if stats else None
        check(lib().gb_bfs_batch(self._h, _ptr(src), src.size, _ptr(levels), C.byref(st) if stats else None))
        return levels, (st.as_dict() if stats else None)

    def bfs_batch_device(self, sources, d_levels, stats=False):
        """Multi-source BFS into a device buffer of k*n int32 (results stay in HBM)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        st = BatchStats()
        check(lib().gb_bfs_batch_device(self._h, _ptr(src), src.size, d_levels.ptr,
                                        C.byref(st) if stats else None))
        return st.as_dict() if stats else None

    def free(self):
        if self._h is not None and self._h.value:
            lib().gb_device_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Continuous-batching BFS query server (gb_server_*).

    Queries (graph id, source) are batched onto `lanes` concurrent GPU streams
    as lanes and the device-memory budget allow; results stay in HBM until
    released.  DeviceGraph objects added to the server are kept alive by it.
    """

    def __init__(self, device=0, max_batch=64, lanes=4, plain=False, memory_budget=0, queue_capacity=0):
        cfg = _ffi.ServerCfg(device, max_batch, lanes, int(plain), memory_budget, queue_capacity)
        h = C.c_void_p()
        check(lib().gb_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._graphs = []

    def add_graph(self, dgraph):
        gid = C.c_int()
        check(lib().gb_server_add_graph(self._h, dgraph._h, C.byref(gid)))
        self._graphs.append(dgraph)
        return gid.value

    def submit(self, gid, source):
        t = C.c_uint64()
        check(lib().gb_server_submit(self._h, int(gid), int(source), C.byref(t)))
        return t.value

    def poll(self, ticket, out=None):
        """Returns (status, latency dict); fills `out` (int32[n]) when done."""
        lat = _ffi.Latency()
        rc = lib().gb_server_poll(self._h, ticket, _ptr(out), C.byref(lat))
        return rc, lat.as_dict()

    def wait(self, ticket, out=None, timeout_ms=-1):
        """Returns (status, latency dict): GB_OK, GB_EAGAIN on timeout, or the
        query's own failure status (GB_ENOMEM, GB_EDEVICE).  Raises for an
        unknown ticket."""
        lat = _ffi.Latency()
        rc = lib().gb_server_wait(self._h, ticket, int(timeout_ms), _ptr(out), C.byref(lat))
        if rc == _ffi.GB_EINVAL:
            check(rc)
        return rc, lat.as_dict()

    def release(self, ticket):
        check(lib().gb_server_release(self._h, ticket))

    def drain(self, timeout_ms=-1):
        rc = lib().gb_server_drain(self._h, int(timeout_ms))
        if rc not in (_ffi.GB_OK, _ffi.GB_EAGAIN):
            check(rc)
        return rc == _ffi.GB_OK

    def stats(self):
        st = _ffi.ServerStats()
        check(lib().gb_server_get_stats(self._h, C.byref(st)))
        return st.as_dict()

    def reset_stats(self):
        check(lib().gb_server_reset_stats(self._h))

    def close(self):
        if self._h is not None and self._h.value:
            lib().gb_server_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def exclusive_scan(values, device=0):
    """Device-wide exclusive int32 scan (gb_exclusive_scan_i32) of a host array."""
    values = np.ascontiguousarray(values, dtype=np.int32)
    d = DeviceArray.from_numpy(device, values)
    check(lib().gb_exclusive_scan_i32(device, d.ptr, d.ptr, values.size))
    return d.numpy()


class csr:
    """Mirror of the reference's ``csr`` class (1D Partition/include/csr.h:9-23).

    ``csr(filename)`` parses the reference edge-list format.  Keyword flags:
    ``one_based`` (MIG/2D loaders), ``symmetrize``, ``keep_order`` (file order
    and duplicates, exactly the reference's layout), ``auto_n`` (grow n to the
    largest id + 1).  ``row_offset`` is int64
    and ``col_idx`` int32.
    """

    def __init__(self, filename=None, one_based=False, symmetrize=False, keep_order=False, auto_n=False,
                 graph=None):
        if graph is None:
            flags = (GB_ONE_BASED if one_based else 0) | (GB_SYMMETRIZE if symmetrize else 0) | \
                    (GB_KEEP_ORDER if keep_order else 0) | (GB_AUTO_N if auto_n else 0)
            graph = Graph.from_file(filename, flags)
        self.graph = graph
        self.row_offset = graph.offsets
        self.col_idx = graph.cols
        self.num_nodes = graph.n
        self.num_edges = graph.m
        self._device_graphs = {}

    def device_graph(self, device=0):
        if device not in self._device_graphs:
            self._device_graphs[device] = DeviceGraph(self.graph, device)
        return self._device_graphs[device]

    def print_info(self):
        print("\n------CSR INFO------")
        print(f"Column indices: {self.num_edges}")
        print(" | ".join(str(int(x)) for x in self.col_idx) + " | ")
        print(f"Row offset: {self.num_nodes}")
        print(" | ".join(str(int(x)) for x in self.row_offset) + " | ")


class gpuBFS:
    """Mirror of the reference's ``gpuBFS`` class (1D Partition/include/gpuBFS.cuh:9-20).

    The constructor runs the whole BFS on one GPU (the graph is uploaded on
    first use and kept resident).  ``host_distance`` holds int32 levels, -1
    for unreached; ``exec_time`` is device time in ms; ``copy_time`` the ms of
    the device-to-host copy of the levels; ``iteration`` is the
    number of BFS levels; ``total_edges_traversed`` the CSR entries scanned.
    """

    def __init__(self, graph, source, device=0, parents=False):
        dg = graph.device_graph(device)
        levels, par, st = dg.bfs(source, parents=parents)
        self.host_distance = levels
        self.parents = par
        self.exec_time = st["ms"]
        self.copy_time = st["copy_ms"]
        self.iteration = st["num_levels"]
        self.total_edges_traversed = st["m_reached"]
        self.vertices_reached = st["n_reached"]

    def print_distance(self, graph=None):
        print("\n------GPU DISTANCE VECTOR------")
        print(" | ".join(str(int(x)) for x in self.host_distance) + " | ")
