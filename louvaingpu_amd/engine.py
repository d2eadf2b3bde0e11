"""Phase-level Python wrapper of the engine's C ABI (``include/louvain_amd.h``).

``Engine`` owns one ``lv_ctx`` (one HIP stream, one HBM arena).  Its methods are
the reference's phases (``LouvainCUDA/LouvainCPP.cpp:129-134``) under their
reference names, plus ``run`` (the whole ``generate_dendogram`` level loop on
the device) and result queries.  All arrays are numpy; nothing here computes
Louvain on the host.
"""
from __future__ import annotations

import ctypes as ct
import weakref

import numpy as np

from . import _lib
from ._lib import LvConfig, LvLevelInfo, LvStats, check

MIN = 1e-7  # the reference's __MIN (LouvainCUDA/LouvainCPP.cpp:55)


def _p32(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_int32))


def _p64(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_int64))


def make_config(gamma=1.0, min_gain=MIN, max_passes=0, max_levels=0, device=0, trace=False,
                timing=False, arena_bytes=0, hub_q=0, hub_part_target=0, chunk_mul=0,
                upload=0, fanout=0) -> LvConfig:
    """lv_config; hub_q / hub_part_target / chunk_mul = 0 take the tuning table;
    upload: 0 encoded from 2^18 edges, 1 plain copies, 2 encoded at any size;
    fanout: 0 extra streams for phases of 2^22 entries or more, 1 never, 2 always."""
    cfg = LvConfig()
    _lib.engine().lv_config_default(ct.byref(cfg))
    cfg.gamma = float(gamma)
    cfg.min_gain = float(min_gain)
    cfg.max_passes = int(max_passes)
    cfg.max_levels = int(max_levels)
    cfg.device = int(device)
    cfg.flags = (_lib.LV_FLAG_TRACE if trace else 0) | (_lib.LV_FLAG_TIMING if timing else 0)
    cfg.arena_bytes = int(arena_bytes)
    cfg.hub_q = int(hub_q)
    cfg.hub_part_target = int(hub_part_target)
    cfg.chunk_mul = int(chunk_mul)
    cfg.upload = int(upload)
    cfg.fanout = int(fanout)
    return cfg


_hip = None


def pinned_empty(count, dtype) -> np.ndarray:
    """An uninitialised numpy array in page-locked host memory (hipHostMalloc),
    freed with hipHostFree when the array is collected.  Copies between it and
    the device run at the link's full rate without a staging copy."""
    global _hip
    if _hip is None:
        _hip = ct.CDLL("libamdhip64.so")
        _hip.hipHostMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t, ct.c_uint]
        _hip.hipHostFree.argtypes = [ct.c_void_p]
    dtype = np.dtype(dtype)
    nbytes = max(1, int(count) * dtype.itemsize)
    p = ct.c_void_p()
    rc = _hip.hipHostMalloc(ct.byref(p), nbytes, 0)
    if rc != 0 or not p.value:
        raise MemoryError("hipHostMalloc of %d bytes failed (%d)" % (nbytes, rc))
    buf = (ct.c_char * nbytes).from_address(p.value)
    weakref.finalize(buf, _hip.hipHostFree, ct.c_void_p(p.value))
    return np.frombuffer(buf, dtype, count=int(count))


def _out(out, n, dtype=np.int32):
    if out is None:
        return np.empty(n, dtype)
    if out.dtype != dtype or len(out) < n or not out.flags.c_contiguous:
        raise ValueError("out must be a contiguous %s array of at least %d entries" % (dtype, n))
    return out[:n]


def weights_array(w):
    """Returns (contiguous array, wtype) for int32 or float32 weights."""
    w = np.ascontiguousarray(w)
    if w.dtype == np.int32:
        return w, _lib.LV_WEIGHT_INT32
    if w.dtype == np.float32:
        return w, _lib.LV_WEIGHT_FLOAT32
    raise TypeError("weights must be int32 or float32, got %s" % w.dtype)


class Engine:
    """One engine context.  Use as a context manager or call ``close()``."""

    def __init__(self, **config):
        self._lib = _lib.engine()
        self.config = make_config(**config)
        h = ct.c_void_p()
        check(self._lib.lv_create(ct.byref(h), ct.byref(self.config)), "lv_create")
        self._h = h
        self.n = 0
        self.e = 0
        self.ids = None

    # lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lv_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # input ----------------------------------------------------------------
    def load_edges(self, n, src, dst, w, ids=None):
        """Graph::setNodes/setEdges + upload (LouvainCUDA/LouvainCPP.cpp:283-287)."""
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        w, wtype = weights_array(w)
        if not (len(src) == len(dst) == len(w)):
            raise ValueError("src, dst and w must have the same length")
        ids_arr = None if ids is None else np.ascontiguousarray(ids, np.int32)
        if ids_arr is not None and len(ids_arr) != n:
            raise ValueError("ids must have n entries")
        check(self._lib.lv_load_edges(self._h, int(n), None if ids_arr is None else _p32(ids_arr),
                                      len(src), _p32(src), _p32(dst),
                                      w.ctypes.data_as(ct.c_void_p), wtype), "lv_load_edges")
        self.n, self.e, self.ids = int(n), len(src), ids_arr
        return self

    # post-processing --------------------------------------------------------
    def cluster_layout(self, comm, ncomm=None):
        """calculateClusterSize + layoutProcess (LouvainCUDA/LouvainCPP.cpp:555-651) on
        the device for a partition given per position: (cluster_end,
        index_in_cluster, x, y)."""
        comm = np.ascontiguousarray(comm, np.int32)
        n = len(comm)
        k = int(comm.max()) + 1 if ncomm is None and n else int(ncomm or 0)
        ce = np.empty(max(k, 1), np.int32)
        ix = np.empty(n, np.int32)
        x = np.empty(n, np.float32)
        y = np.empty(n, np.float32)
        fp = ct.POINTER(ct.c_float)
        check(self._lib.lv_cluster_layout(self._h, n, _p32(comm), k, _p32(ce), _p32(ix),
                                          x.ctypes.data_as(fp), y.ctypes.data_as(fp)),
              "lv_cluster_layout")
        return ce[:k], ix, x, y

    # phases (reference names) ---------------------------------------------
    def CalculateNeighbour(self):
        check(self._lib.lv_neighbour_compute(self._h), "lv_neighbour_compute")

    def initCUDA(self):
        check(self._lib.lv_init(self._h), "lv_init")

    def modularityCUDA(self) -> float:
        q = ct.c_double()
        check(self._lib.lv_modularity(self._h, ct.byref(q)), "lv_modularity")
        return q.value

    def oneLevelCUDA(self, lower: float):
        q = ct.c_double()
        p = ct.c_int32()
        check(self._lib.lv_one_level(self._h, float(lower), ct.byref(q), ct.byref(p)),
              "lv_one_level")
        return q.value, p.value

    def renumberCUDA(self) -> int:
        k = ct.c_int32()
        check(self._lib.lv_renumber(self._h, ct.byref(k)), "lv_renumber")
        return k.value

    def inducedGraphCUDA(self):
        check(self._lib.lv_induce(self._h), "lv_induce")

    # whole level loop -----------------------------------------------------
    def run(self) -> int:
        lv = ct.c_int32()
        check(self._lib.lv_run(self._h, ct.byref(lv)), "lv_run")
        return lv.value

    def run_partitions(self, node_bufs, comm_bufs, final):
        """run() with the dendrogram delivered to the host during the run:
        Partition of level l into node_bufs[l] / comm_bufs[l] (int32 arrays of at
        least n entries each; page-locked ones let the copies overlap the run),
        the final partition into `final`.  Returns the number of levels."""
        k = len(node_bufs)
        if len(comm_bufs) != k:
            raise ValueError("node_bufs and comm_bufs must pair up")
        for a in list(node_bufs) + list(comm_bufs) + [final]:
            if a.dtype != np.int32 or len(a) < self.n or not a.flags.c_contiguous:
                raise ValueError("buffers must be contiguous int32 arrays of at least n entries")
        np_ = (ct.c_void_p * max(k, 1))(*[a.ctypes.data for a in node_bufs])
        cp_ = (ct.c_void_p * max(k, 1))(*[a.ctypes.data for a in comm_bufs])
        lv = ct.c_int32()
        check(self._lib.lv_run_partitions(self._h, k, np_, cp_, _p32(final), ct.byref(lv)),
              "lv_run_partitions")
        return lv.value

    # results --------------------------------------------------------------
    def num_levels(self) -> int:
        lv = ct.c_int32()
        check(self._lib.lv_num_levels(self._h, ct.byref(lv)), "lv_num_levels")
        return lv.value

    def level_info(self, level) -> LvLevelInfo:
        info = LvLevelInfo()
        check(self._lib.lv_level_info_get(self._h, int(level), ct.byref(info)), "lv_level_info_get")
        return info

    def level_partition(self, level, out=None):
        """(node, comm) of the level's Partition; out=(node, comm) buffers to fill."""
        n = self.level_info(level).n
        node = _out(None if out is None else out[0], n)
        comm = _out(None if out is None else out[1], n)
        check(self._lib.lv_level_partition(self._h, int(level), _p32(node), _p32(comm)),
              "lv_level_partition")
        return node, comm

    def level_assignment(self, level):
        n = self.level_info(level).n
        out = np.empty(n, np.int32)
        check(self._lib.lv_level_assignment(self._h, int(level), _p32(out)), "lv_level_assignment")
        return out

    def flatten(self, level):
        node = np.empty(self.n, np.int32)
        comm = np.empty(self.n, np.int32)
        check(self._lib.lv_flatten(self._h, int(level), _p32(node), _p32(comm)), "lv_flatten")
        return node, comm

    def final_assignment(self, out=None):
        out = _out(out, self.n)
        check(self._lib.lv_final_assignment(self._h, _p32(out)), "lv_final_assignment")
        return out

    def final_modularity(self) -> float:
        q = ct.c_double()
        check(self._lib.lv_final_modularity(self._h, ct.byref(q)), "lv_final_modularity")
        return q.value

    def trace_next(self, level, p):
        n = self.level_info(level).n
        out = np.empty(n, np.int32)
        check(self._lib.lv_trace_next(self._h, int(level), int(p), _p32(out)), "lv_trace_next")
        return out

    def trace_pass_q(self, level, p) -> float:
        q = ct.c_double()
        check(self._lib.lv_trace_pass_q(self._h, int(level), int(p), ct.byref(q)), "lv_trace_pass_q")
        return q.value

    def current_level(self):
        lv, n, nnz = ct.c_int32(), ct.c_int32(), ct.c_int64()
        check(self._lib.lv_current_level(self._h, ct.byref(lv), ct.byref(n), ct.byref(nnz)),
              "lv_current_level")
        return lv.value, n.value, nnz.value

    def level_csr(self):
        _, n, nnz = self.current_level()
        rp = np.empty(n + 1, np.int64)
        col = np.empty(max(nnz, 1), np.int32)
        w = np.empty(max(nnz, 1), np.int64)
        check(self._lib.lv_level_csr(self._h, _p64(rp), _p32(col), _p64(w)), "lv_level_csr")
        return rp, col[:nnz], w[:nnz]

    def level_vertex(self):
        _, n, _ = self.current_level()
        k = np.empty(n, np.int64)
        loop = np.empty(n, np.int64)
        selfsum = np.empty(n, np.int64)
        check(self._lib.lv_level_vertex(self._h, _p64(k), _p64(loop), _p64(selfsum)),
              "lv_level_vertex")
        return k, loop, selfsum

    def stats(self) -> LvStats:
        s = LvStats()
        check(self._lib.lv_get_stats(self._h, ct.byref(s)), "lv_get_stats")
        return s

    def reset_stats(self):
        check(self._lib.lv_reset_stats(self._h), "lv_reset_stats")

    def last_run_ms(self) -> float:
        ms = ct.c_double()
        check(self._lib.lv_last_run_ms(self._h, ct.byref(ms)), "lv_last_run_ms")
        return ms.value


def generate_dendogram_phases(eng: Engine, min_gain=MIN, max_levels=0):
    """The reference's level loop (LouvainCUDA/LouvainCPP.cpp:334-507) driven phase
    by phase through the C ABI; returns the per-level (node, community) arrays."""
    dendro = []
    eng.CalculateNeighbour()
    eng.initCUDA()
    mod = eng.modularityCUDA()
    new_mod, _ = eng.oneLevelCUDA(mod)
    eng.renumberCUDA()
    dendro.append(eng.level_partition(0))
    mod = new_mod
    eng.inducedGraphCUDA()
    level = 1
    while True:
        eng.CalculateNeighbour()
        eng.initCUDA()
        new_mod, _ = eng.oneLevelCUDA(mod)
        eng.renumberCUDA()
        dendro.append(eng.level_partition(level))
        if new_mod - mod < min_gain:
            break
        if max_levels and level + 1 >= max_levels:
            break
        mod = new_mod
        eng.inducedGraphCUDA()
        level += 1
    return dendro
