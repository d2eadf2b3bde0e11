"""ctypes bindings of liblouvain_amd.so (C ABI in include/louvain_amd.h) and
liblouvain_host.so (generators, generate_dendogram).

The libraries are built in-tree by ``louvaingpu_amd.build`` (``make`` in
``csrc/``).  Loading fails loudly when they are missing: there is no fallback
path.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ENGINE_PATH = os.path.join(HERE, "liblouvain_amd.so")
HOST_PATH = os.path.join(HERE, "liblouvain_host.so")

LV_OK = 0
LV_ERR_NOMEM = -1
LV_ERR_INVALID_INPUT = -2
LV_ERR_RANGE = -3
LV_ERR_ARG = -4
LV_ERR_DEVICE = -5
LV_ERR_NO_TRACE = -6
LV_ERR_TOO_LARGE = -7

LV_STATIC_SCRATCH_BYTES = 256 << 20

LV_WEIGHT_INT32 = 0
LV_WEIGHT_FLOAT32 = 1

LV_FLAG_TRACE = 1
LV_FLAG_TIMING = 2

LV_SWEEP_TIERS = 12

# Every function the C header declares (checked by the test suite).
ENGINE_SYMBOLS = (
    "lv_config_default", "lv_strerror", "lv_create", "lv_destroy", "lv_load_edges",
    "lv_neighbour_compute", "lv_init", "lv_modularity", "lv_one_level", "lv_renumber",
    "lv_induce", "lv_run", "lv_num_levels", "lv_level_info_get", "lv_level_partition",
    "lv_level_assignment", "lv_flatten", "lv_final_assignment", "lv_final_modularity",
    "lv_cluster_layout", "lv_trace_next", "lv_trace_pass_q", "lv_level_csr", "lv_level_vertex", "lv_current_level",
    "lv_get_stats", "lv_reset_stats", "lv_last_run_ms",
    "lv_batch_load", "lv_batch_load_graphs", "lv_batch_run", "lv_batch_num_components", "lv_batch_num_levels",
    "lv_batch_level_info", "lv_batch_level_assignment", "lv_batch_final_assignment",
    "lv_run_static", "lv_run_partitions",
)
HOST_SYMBOLS = (
    "lvh_generate_dendrogram", "lvh_num_levels", "lvh_level_count", "lvh_level_partition",
    "lvh_partition_at_level", "lvh_get_partitions", "lvh_count_partitions", "lvh_free",
    "lvh_cluster_layout",
    "lvg_sbm", "lvg_rmat", "lvg_num_nodes", "lvg_num_edges", "lvg_wtype", "lvg_copy", "lvg_free",
    "lvh_load_text", "lvh_file_num_nodes", "lvh_file_num_edges", "lvh_file_copy", "lvh_file_free",
    "lvh_binary_header", "lvh_binary_read", "lvh_binary_write",
)
SERVE_SYMBOLS = (
    "lvs_create", "lvs_submit", "lvs_wait", "lvs_get_stats", "lvs_now_ms", "lvs_destroy",
)


class LvConfig(ct.Structure):
    _fields_ = [
        ("gamma", ct.c_double),
        ("min_gain", ct.c_double),
        ("max_passes", ct.c_int32),
        ("max_levels", ct.c_int32),
        ("device", ct.c_int32),
        ("flags", ct.c_int32),
        ("arena_bytes", ct.c_uint64),
        ("hub_q", ct.c_int32),
        ("hub_part_target", ct.c_int32),
        ("chunk_mul", ct.c_int32),
        ("upload", ct.c_int32),
        ("fanout", ct.c_int32),
    ]


class LvLevelInfo(ct.Structure):
    _fields_ = [
        ("n", ct.c_int32),
        ("k", ct.c_int32),
        ("nnz", ct.c_int64),
        ("m_units", ct.c_int64),
        ("passes", ct.c_int32),
        ("scale", ct.c_int32),
        ("q", ct.c_double),
        ("lower", ct.c_double),
    ]


class LvsResult(ct.Structure):
    _fields_ = [
        ("status", ct.c_int32),
        ("levels", ct.c_int32),
        ("communities", ct.c_int32),
        ("batch_size", ct.c_int32),
        ("q", ct.c_double),
        ("t_submit_ms", ct.c_double),
        ("t_start_ms", ct.c_double),
        ("t_done_ms", ct.c_double),
    ]


class LvsStats(ct.Structure):
    _fields_ = [
        ("submitted", ct.c_int64),
        ("completed", ct.c_int64),
        ("sets", ct.c_int64),
        ("max_set", ct.c_int32),
        ("workers", ct.c_int32),
    ]


class LvStats(ct.Structure):
    _fields_ = [
        ("csr_ms", ct.c_double),
        ("init_ms", ct.c_double),
        ("sweep_ms", ct.c_double),
        ("apply_ms", ct.c_double),
        ("renumber_ms", ct.c_double),
        ("coarsen_ms", ct.c_double),
        ("flatten_ms", ct.c_double),
        ("tier_ms", ct.c_double * LV_SWEEP_TIERS),
        ("tier_launches", ct.c_int64 * LV_SWEEP_TIERS),
        ("tier_bytes", ct.c_double * LV_SWEEP_TIERS),
        ("tier_cand", ct.c_double * LV_SWEEP_TIERS),
        ("sweep_bytes", ct.c_double),
        ("sweep_entries", ct.c_int64),
        ("levels", ct.c_int32),
        ("passes", ct.c_int32),
        ("arena_capacity", ct.c_uint64),
        ("arena_high_water", ct.c_uint64),
        ("upload_bytes", ct.c_uint64),
    ]


_I32P = ct.POINTER(ct.c_int32)
_I64P = ct.POINTER(ct.c_int64)
_DP = ct.POINTER(ct.c_double)
_VP = ct.c_void_p

_engine = None
_host = None


class NativeLibraryMissing(ImportError):
    pass


def _open(path):
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            "%s is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C louvaingpu_amd/csrc`" % path)
    return ct.CDLL(path, mode=ct.RTLD_GLOBAL)


def engine():
    """The loaded liblouvain_amd.so with argtypes set."""
    global _engine
    if _engine is not None:
        return _engine
    lib = _open(ENGINE_PATH)
    ctx = ct.c_void_p
    sig = {
        "lv_config_default": (None, [ct.POINTER(LvConfig)]),
        "lv_strerror": (ct.c_char_p, [ct.c_int]),
        "lv_create": (ct.c_int, [ct.POINTER(ctx), ct.POINTER(LvConfig)]),
        "lv_destroy": (ct.c_int, [ctx]),
        "lv_load_edges": (ct.c_int, [ctx, ct.c_int32, _I32P, ct.c_int64, _I32P, _I32P, _VP,
                                     ct.c_int]),
        "lv_neighbour_compute": (ct.c_int, [ctx]),
        "lv_init": (ct.c_int, [ctx]),
        "lv_modularity": (ct.c_int, [ctx, _DP]),
        "lv_one_level": (ct.c_int, [ctx, ct.c_double, _DP, _I32P]),
        "lv_renumber": (ct.c_int, [ctx, _I32P]),
        "lv_induce": (ct.c_int, [ctx]),
        "lv_run": (ct.c_int, [ctx, _I32P]),
        "lv_run_partitions": (ct.c_int, [ctx, ct.c_int32, ct.POINTER(ct.c_void_p),
                                         ct.POINTER(ct.c_void_p), _I32P, _I32P]),
        "lv_num_levels": (ct.c_int, [ctx, _I32P]),
        "lv_level_info_get": (ct.c_int, [ctx, ct.c_int32, ct.POINTER(LvLevelInfo)]),
        "lv_level_partition": (ct.c_int, [ctx, ct.c_int32, _I32P, _I32P]),
        "lv_level_assignment": (ct.c_int, [ctx, ct.c_int32, _I32P]),
        "lv_flatten": (ct.c_int, [ctx, ct.c_int32, _I32P, _I32P]),
        "lv_final_assignment": (ct.c_int, [ctx, _I32P]),
        "lv_final_modularity": (ct.c_int, [ctx, _DP]),
        "lv_cluster_layout": (ct.c_int, [ctx, ct.c_int32, _I32P, ct.c_int32, _I32P, _I32P,
                                         ct.POINTER(ct.c_float), ct.POINTER(ct.c_float)]),
        "lv_trace_next": (ct.c_int, [ctx, ct.c_int32, ct.c_int32, _I32P]),
        "lv_trace_pass_q": (ct.c_int, [ctx, ct.c_int32, ct.c_int32, _DP]),
        "lv_level_csr": (ct.c_int, [ctx, _I64P, _I32P, _I64P]),
        "lv_level_vertex": (ct.c_int, [ctx, _I64P, _I64P, _I64P]),
        "lv_current_level": (ct.c_int, [ctx, _I32P, _I32P, _I64P]),
        "lv_get_stats": (ct.c_int, [ctx, ct.POINTER(LvStats)]),
        "lv_reset_stats": (ct.c_int, [ctx]),
        "lv_last_run_ms": (ct.c_int, [ctx, _DP]),
        "lv_batch_load": (ct.c_int, [ctx, ct.c_int32, _I32P, _I64P, _I32P, _I32P, _VP, ct.c_int,
                                     ct.c_int32, _DP]),
        "lv_batch_load_graphs": (ct.c_int, [ctx, ct.c_int32, _I32P, _I64P, ct.POINTER(ct.c_void_p),
                                            ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_void_p),
                                            ct.c_int, ct.c_int32, _DP]),
        "lv_batch_run": (ct.c_int, [ctx]),
        "lv_batch_num_components": (ct.c_int, [ctx, _I32P]),
        "lv_batch_num_levels": (ct.c_int, [ctx, ct.c_int32, _I32P]),
        "lv_batch_level_info": (ct.c_int, [ctx, ct.c_int32, ct.c_int32, ct.POINTER(LvLevelInfo)]),
        "lv_batch_level_assignment": (ct.c_int, [ctx, ct.c_int32, ct.c_int32, _I32P]),
        "lv_batch_final_assignment": (ct.c_int, [ctx, ct.c_int32, _I32P]),
        "lv_run_static": (ct.c_int, [ct.POINTER(LvConfig), ct.c_void_p, ct.c_int32, _I32P,
                                     ct.c_int64, _I32P, _I32P, _VP, ct.c_int, _I32P, _I32P, _DP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _engine = lib
    return lib


def host():
    """The loaded liblouvain_host.so with argtypes set."""
    global _host
    if _host is not None:
        return _host
    engine()
    lib = _open(HOST_PATH)
    d = ct.c_void_p
    sig = {
        "lvh_generate_dendrogram": (ct.c_int, [ct.c_int32, _I32P, ct.c_int64, _I32P, _I32P,
                                               ct.POINTER(ct.c_float), ct.POINTER(LvConfig),
                                               ct.POINTER(d)]),
        "lvh_num_levels": (ct.c_int, [d]),
        "lvh_level_count": (ct.c_int32, [d, ct.c_int]),
        "lvh_level_partition": (ct.c_int, [d, ct.c_int, _I32P, _I32P]),
        "lvh_partition_at_level": (ct.c_int, [d, ct.c_int, _I32P, _I32P]),
        "lvh_get_partitions": (ct.c_int, [d, ct.c_int, _I32P, _I32P]),
        "lvh_count_partitions": (ct.c_int, [d, ct.c_int]),
        "lvh_free": (None, [d]),
        "lvh_cluster_layout": (ct.c_int, [ct.c_int32, _I32P, ct.c_int32, _I32P, _I32P, ct.c_int32,
                                          _I32P, _I32P, _I32P, ct.POINTER(ct.c_float),
                                          ct.POINTER(ct.c_float)]),
        "lvg_sbm": (ct.c_void_p, [ct.c_int32, ct.c_int32, ct.c_double, ct.c_double, ct.c_uint64]),
        "lvg_rmat": (ct.c_void_p, [ct.c_int32, ct.c_int64, ct.c_double, ct.c_double,
                                   ct.c_double, ct.c_uint64, ct.c_int]),
        "lvg_num_nodes": (ct.c_int32, [d]),
        "lvg_num_edges": (ct.c_int64, [d]),
        "lvg_wtype": (ct.c_int, [d]),
        "lvg_copy": (ct.c_int, [d, _I32P, _I32P, ct.c_void_p]),
        "lvg_free": (None, [d]),
        "lvh_load_text": (ct.c_int, [ct.c_char_p, ct.c_char_p, ct.POINTER(d)]),
        "lvh_file_num_nodes": (ct.c_int32, [d]),
        "lvh_file_num_edges": (ct.c_int64, [d]),
        "lvh_file_copy": (ct.c_int, [d, _I32P, _I32P, _I32P, ct.POINTER(ct.c_float)]),
        "lvh_file_free": (None, [d]),
        "lvh_binary_header": (ct.c_int, [ct.c_char_p, _I32P, _I64P, ct.POINTER(ct.c_int),
                                         ct.POINTER(ct.c_int)]),
        "lvh_binary_read": (ct.c_int, [ct.c_char_p, _I32P, _I32P, _I32P, ct.c_void_p]),
        "lvh_binary_write": (ct.c_int, [ct.c_char_p, ct.c_int32, _I32P, ct.c_int64, _I32P, _I32P,
                                        ct.c_void_p, ct.c_int]),
        "lvs_create": (ct.c_int, [ct.POINTER(d), ct.POINTER(LvConfig), ct.c_int32, ct.c_uint64]),
        "lvs_submit": (ct.c_int, [d, ct.c_int32, ct.c_int64, _I32P, _I32P, ct.c_void_p, ct.c_int,
                                  ct.c_double, _I64P]),
        "lvs_wait": (ct.c_int, [d, ct.c_int64, _I32P, ct.POINTER(LvsResult)]),
        "lvs_get_stats": (ct.c_int, [d, ct.POINTER(LvsStats)]),
        "lvs_now_ms": (ct.c_double, [d]),
        "lvs_destroy": (ct.c_int, [d]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _host = lib
    return lib


def strerror(status: int) -> str:
    return engine().lv_strerror(status).decode()


class LouvainError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        super().__init__("%s: %s (%d)" % (where, strerror(status), status))


def check(status, where=""):
    if status != LV_OK:
        raise LouvainError(status, where)
    return status
