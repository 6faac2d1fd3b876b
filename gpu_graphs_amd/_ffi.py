"""ctypes binding of libgpugraphs.so (the C ABI in include/gpugraphs/gb.h).

The shared library is built in-tree by gpu_graphs_amd/build.py.  There is no
fallback: if the library is missing, importing the binding raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpugraphs.so")

GB_OK = 0
GB_EINVAL = -1
GB_EIO = -2
GB_ENOMEM = -3
GB_EDEVICE = -4
GB_ERANGE = -5
GB_EBUSY = -6
GB_EAGAIN = -7

GB_ONE_BASED = 1
GB_SYMMETRIZE = 2
GB_KEEP_ORDER = 4
GB_DROP_SELF_LOOPS = 8
GB_AUTO_N = 16

GB_MODE_ASYNC = 0
GB_MODE_LEVEL_LOOP = 1
GB_MODE_GRAPH = 2

GB_DEEP_AUTO = 0
GB_DEEP_ON = 1
GB_DEEP_OFF = 2
GB_TAIL_AUTO = -1
GB_TAIL_OFF = -2
GB_TAIL_FINISHED = 1
GB_TAIL_DECLINED = 2
GB_TAIL_HANDED_BACK = 3

_NAMES = {
    GB_EINVAL: "GB_EINVAL", GB_EIO: "GB_EIO", GB_ENOMEM: "GB_ENOMEM", GB_EDEVICE: "GB_EDEVICE",
    GB_ERANGE: "GB_ERANGE", GB_EBUSY: "GB_EBUSY", GB_EAGAIN: "GB_EAGAIN",
}


class GBError(RuntimeError):
    """A GB_E* status returned by the library, with gb_last_error()'s text."""

    def __init__(self, code, message):
        super().__init__(f"{_NAMES.get(code, code)}: {message}")
        self.code = code


class Stats(C.Structure):
    _fields_ = [
        ("ms", C.c_double),
        ("kernel_ms", C.c_double),
        ("num_levels", C.c_int32),
        ("reserved", C.c_int32),
        ("n_reached", C.c_int64),
        ("m_reached", C.c_int64),
        ("copy_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class BatchStats(C.Structure):
    _fields_ = [
        ("ms", C.c_double),
        ("num_levels", C.c_int32),
        ("k", C.c_int32),
        ("n_reached", C.c_int64 * 64),
        ("m_reached", C.c_int64 * 64),
    ]

    def as_dict(self):
        k = self.k
        return {"ms": self.ms, "num_levels": self.num_levels, "k": k,
                "n_reached": list(self.n_reached)[:k], "m_reached": list(self.m_reached)[:k]}


class ServerCfg(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("max_batch", C.c_int),
        ("lanes", C.c_int),
        ("plain", C.c_int),
        ("memory_budget", C.c_int64),
        ("queue_capacity", C.c_int64),
    ]


class Latency(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("run_ms", C.c_double), ("total_ms", C.c_double),
                ("batch_size", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64), ("rejected", C.c_int64),
        ("batches", C.c_int64), ("mean_batch", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
        ("max_ms", C.c_double), ("qps", C.c_double), ("memory_used", C.c_int64), ("memory_budget", C.c_int64),
        ("workspace_drops", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneResult(C.Structure):
    _fields_ = [
        ("default_ms", C.c_double), ("best_ms", C.c_double), ("tried", C.c_int32), ("block", C.c_int32),
        ("edges_per_thread", C.c_int32), ("alpha", C.c_int32), ("beta", C.c_int32), ("low_degree", C.c_int32),
        ("frontier_edges", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)

# name -> (restype, argtypes).  Mirrors include/gpugraphs/gb.h.
SIGNATURES = {
    "gb_last_error": (C.c_char_p, []),
    "gb_version": (C.c_char_p, []),
    "gb_graph_load_edgelist": (C.c_int, [C.c_char_p, C.c_uint32, PP]),
    "gb_graph_from_edges": (C.c_int, [C.c_int64, C.c_int64, P, P, C.c_uint32, PP]),
    "gb_graph_from_csr": (C.c_int, [C.c_int64, P, P, C.c_uint32, PP]),
    "gb_graph_generate_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                         C.c_uint64, C.c_int, PP]),
    "gb_graph_generate_er": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, C.c_uint64, PP]),
    "gb_graph_generate_grid": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_uint64, PP]),
    "gb_graph_info": (C.c_int, [P, i64p, i64p]),
    "gb_graph_csr": (C.c_int, [P, C.POINTER(P), C.POINTER(P)]),
    "gb_graph_degree_buckets": (C.c_int, [P, i64p, C.POINTER(P)]),
    "gb_graph_is_symmetric": (C.c_int, [P, C.POINTER(C.c_int)]),
    "gb_graph_write_edgelist": (C.c_int, [P, C.c_char_p]),
    "gb_graph_save_binary": (C.c_int, [P, C.c_char_p]),
    "gb_graph_load_binary": (C.c_int, [C.c_char_p, PP]),
    "gb_graph_free": (None, [P]),
    "gb_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "gb_device_upload": (C.c_int, [P, C.c_int, PP]),
    "gb_dgraph_info": (C.c_int, [P, i64p, i64p, C.POINTER(C.c_int)]),
    "gb_dgraph_memory": (C.c_int, [P, i64p]),
    "gb_device_free": (None, [P]),
    "gb_bfs": (C.c_int, [P, C.c_int32, P, P, C.POINTER(Stats)]),
    "gb_bfs_device": (C.c_int, [P, C.c_int32, P, P, C.POINTER(Stats)]),
    "gb_bfs_config": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64]),
    "gb_bfs_get_config": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), i64p]),
    "gb_bfs_direction_config": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
    "gb_bfs_get_direction_config": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                              C.POINTER(C.c_int)]),
    "gb_bfs_last_bottom_up": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gb_bfs_launch_hints": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gb_bfs_batch": (C.c_int, [P, P, C.c_int, P, C.POINTER(BatchStats)]),
    "gb_bfs_batch_device": (C.c_int, [P, P, C.c_int, P, C.POINTER(BatchStats)]),
    "gb_batch_config": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int64]),
    "gb_batch_get_config": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int64)]),
    "gb_batch_direction_config": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
    "gb_batch_get_direction_config": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gb_batch_last_bottom_up": (C.c_int, [P, C.POINTER(C.c_int)]),
    "gb_bfs_deep_config": (C.c_int, [P, C.c_int]),
    "gb_batch_deep_config": (C.c_int, [P, C.c_int]),
    "gb_batch_get_deep_config": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gb_bfs_get_deep_config": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gb_bfs_tail_config": (C.c_int, [P, C.c_int]),
    "gb_bfs_get_tail_config": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gb_bfs_autotune": (C.c_int, [P, P, C.c_int, C.c_int, C.POINTER(TuneResult)]),
    "gb_batch_autotune": (C.c_int, [P, P, C.c_int, C.c_int, C.POINTER(TuneResult)]),
    "gb_bfs_apply_tuning": (C.c_int, [P, C.POINTER(TuneResult)]),
    "gb_batch_apply_tuning": (C.c_int, [P, C.POINTER(TuneResult)]),
    "gb_server_create": (C.c_int, [C.POINTER(ServerCfg), PP]),
    "gb_server_add_graph": (C.c_int, [P, P, C.POINTER(C.c_int)]),
    "gb_server_submit": (C.c_int, [P, C.c_int, C.c_int32, C.POINTER(C.c_uint64)]),
    "gb_server_poll": (C.c_int, [P, C.c_uint64, P, C.POINTER(Latency)]),
    "gb_server_wait": (C.c_int, [P, C.c_uint64, C.c_int, P, C.POINTER(Latency)]),
    "gb_server_result_device": (C.c_int, [P, C.c_uint64, C.POINTER(P)]),
    "gb_server_release": (C.c_int, [P, C.c_uint64]),
    "gb_server_drain": (C.c_int, [P, C.c_int]),
    "gb_server_get_stats": (C.c_int, [P, C.POINTER(ServerStats)]),
    "gb_server_reset_stats": (C.c_int, [P]),
    "gb_server_free": (None, [P]),
    "gb_exclusive_scan_i32": (C.c_int, [C.c_int, P, P, C.c_int64]),
    "gb_exclusive_scan_i32_noalloc": (C.c_int, [C.c_int, P, P, C.c_int64]),
    "gb_exclusive_scan_noalloc_max": (C.c_int64, []),
    "gb_dmalloc": (C.c_int, [C.c_int, C.c_size_t, PP]),
    "gb_dfree": (C.c_int, [C.c_int, P]),
    "gb_memcpy_htod": (C.c_int, [C.c_int, P, P, C.c_size_t]),
    "gb_memcpy_dtoh": (C.c_int, [C.c_int, P, P, C.c_size_t]),
    "gb_device_synchronize": (C.c_int, [C.c_int]),
}

_lib = None


def lib():
    """Loads libgpugraphs.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m gpu_graphs_amd.build` "
                "(there is no CPU fallback)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(status):
    if status != GB_OK:
        msg = lib().gb_last_error()
        raise GBError(status, msg.decode() if msg else "")
    return status
