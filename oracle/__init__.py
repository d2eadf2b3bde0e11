"""ctypes wrapper around the Louvain oracle (``louvain_oracle.c``).

TEST INFRASTRUCTURE ONLY: the test suite, the smoke check and the CPU-baseline
leg of ``bench.py`` use it as the checker.  The engine never imports it.

``run(...)`` executes the reference's full multi-level Louvain on the CPU and
returns an :class:`OracleRun` exposing every level (CSR, per-pass ``next[]``,
renumbered assignment, Partition arrays, Q values) and the flattened result.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

WTYPE_INT32 = 0
WTYPE_FLOAT32 = 1

_I32P = ct.POINTER(ct.c_int32)
_I64P = ct.POINTER(ct.c_int64)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ct.CDLL(_LIB_PATH)
    lib.orc_run_louvain.restype = ct.c_void_p
    lib.orc_run_louvain.argtypes = [ct.c_int32, _I32P, ct.c_int64, _I32P, _I32P, ct.c_void_p,
                                    ct.c_int, ct.c_double, ct.c_double, ct.c_int32, ct.c_int,
                                    ct.c_int]
    for name in ("orc_status", "orc_scale", "orc_num_levels"):
        getattr(lib, name).restype = ct.c_int
        getattr(lib, name).argtypes = [ct.c_void_p]
    for name, rt in (("orc_level_n", ct.c_int32), ("orc_level_e", ct.c_int64),
                     ("orc_level_m", ct.c_int64), ("orc_level_k", ct.c_int32),
                     ("orc_level_passes", ct.c_int32), ("orc_level_q", ct.c_double),
                     ("orc_level_q_init", ct.c_double), ("orc_level_lower", ct.c_double)):
        getattr(lib, name).restype = rt
        getattr(lib, name).argtypes = [ct.c_void_p, ct.c_int]
    lib.orc_level_pass_q.restype = ct.c_double
    lib.orc_level_pass_q.argtypes = [ct.c_void_p, ct.c_int, ct.c_int]
    lib.orc_level_pass_next.restype = ct.c_int
    lib.orc_level_pass_next.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, _I32P]
    for name in ("orc_level_newc", "orc_level_cfinal"):
        getattr(lib, name).restype = ct.c_int
        getattr(lib, name).argtypes = [ct.c_void_p, ct.c_int, _I32P]
    lib.orc_level_partition.restype = ct.c_int
    lib.orc_level_partition.argtypes = [ct.c_void_p, ct.c_int, _I32P, _I32P]
    lib.orc_level_edges.restype = ct.c_int
    lib.orc_level_edges.argtypes = [ct.c_void_p, ct.c_int, _I32P, _I32P, _I64P]
    lib.orc_level_csr.restype = ct.c_int
    lib.orc_level_csr.argtypes = [ct.c_void_p, ct.c_int, _I64P, _I32P, _I64P]
    lib.orc_level_vertex.restype = ct.c_int
    lib.orc_level_vertex.argtypes = [ct.c_void_p, ct.c_int, _I64P, _I64P]
    lib.orc_flatten.restype = ct.c_int
    lib.orc_flatten.argtypes = [ct.c_void_p, ct.c_int, _I32P, _I32P]
    lib.orc_final_comm.restype = ct.c_int
    lib.orc_final_comm.argtypes = [ct.c_void_p, _I32P]
    lib.orc_modularity_fp64.restype = ct.c_double
    lib.orc_modularity_fp64.argtypes = [ct.c_int32, _I32P, ct.c_int64, _I32P, _I32P, ct.c_void_p,
                                        ct.c_int, _I32P, ct.c_double]
    lib.orc_final_modularity.restype = ct.c_double
    lib.orc_final_modularity.argtypes = [ct.c_void_p, _I32P, _I32P, ct.c_void_p, ct.c_int]
    lib.orc_free.restype = None
    lib.orc_free.argtypes = [ct.c_void_p]
    _lib = lib
    return lib


def _p32(a):
    return a.ctypes.data_as(_I32P)


def _p64(a):
    return a.ctypes.data_as(_I64P)


def _weights(w):
    w = np.ascontiguousarray(w)
    if w.dtype == np.float32:
        return w, WTYPE_FLOAT32
    if w.dtype == np.int32:
        return w, WTYPE_INT32
    raise TypeError("weights must be int32 or float32, got %s" % w.dtype)


class OracleError(RuntimeError):
    pass


class OracleRun:
    """Result of one oracle run; frees the native state on close/GC."""

    def __init__(self, handle, n, src, dst, w, wtype, ids):
        self._h = handle
        self._lib = _load()
        self.n = n
        self._src, self._dst, self._w, self._wtype, self._ids = src, dst, w, wtype, ids
        st = self._lib.orc_status(handle)
        if st != 0:
            self.close()
            raise OracleError("oracle status %d" % st)
        self.scale = self._lib.orc_scale(handle)
        self.num_levels = self._lib.orc_num_levels(handle)

    def close(self):
        if self._h:
            self._lib.orc_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # per level -----------------------------------------------------------
    def level_n(self, l):
        return self._lib.orc_level_n(self._h, l)

    def level_e(self, l):
        return self._lib.orc_level_e(self._h, l)

    def level_m(self, l):
        return self._lib.orc_level_m(self._h, l)

    def level_k(self, l):
        return self._lib.orc_level_k(self._h, l)

    def level_passes(self, l):
        return self._lib.orc_level_passes(self._h, l)

    def level_q(self, l):
        return self._lib.orc_level_q(self._h, l)

    def level_q_init(self, l):
        return self._lib.orc_level_q_init(self._h, l)

    def level_lower(self, l):
        return self._lib.orc_level_lower(self._h, l)

    def pass_q(self, l, p):
        return self._lib.orc_level_pass_q(self._h, l, p)

    def pass_next(self, l, p):
        out = np.empty(self.level_n(l), np.int32)
        if self._lib.orc_level_pass_next(self._h, l, p, _p32(out)) != 0:
            raise OracleError("no trace for level %d pass %d" % (l, p))
        return out

    def newc(self, l):
        out = np.empty(self.level_n(l), np.int32)
        self._lib.orc_level_newc(self._h, l, _p32(out))
        return out

    def cfinal(self, l):
        out = np.empty(self.level_n(l), np.int32)
        self._lib.orc_level_cfinal(self._h, l, _p32(out))
        return out

    def partition(self, l):
        n = self.level_n(l)
        node = np.empty(n, np.int32)
        comm = np.empty(n, np.int32)
        self._lib.orc_level_partition(self._h, l, _p32(node), _p32(comm))
        return node, comm

    def edges(self, l):
        e = self.level_e(l)
        s = np.empty(e, np.int32)
        t = np.empty(e, np.int32)
        w = np.empty(e, np.int64)
        self._lib.orc_level_edges(self._h, l, _p32(s), _p32(t), _p64(w))
        return s, t, w

    def csr(self, l):
        n, e = self.level_n(l), self.level_e(l)
        rp = np.empty(n + 1, np.int64)
        col = np.empty(2 * e, np.int32)
        w = np.empty(2 * e, np.int64)
        self._lib.orc_level_csr(self._h, l, _p64(rp), _p32(col), _p64(w))
        return rp, col, w

    def vertex(self, l):
        n = self.level_n(l)
        k = np.empty(n, np.int64)
        loop = np.empty(n, np.int64)
        self._lib.orc_level_vertex(self._h, l, _p64(k), _p64(loop))
        return k, loop

    # whole run -----------------------------------------------------------
    def flatten(self, level):
        node = np.empty(self.n, np.int32)
        comm = np.empty(self.n, np.int32)
        if self._lib.orc_flatten(self._h, level, _p32(node), _p32(comm)) != 0:
            raise OracleError("bad level %d" % level)
        return node, comm

    def final_comm(self):
        out = np.empty(self.n, np.int32)
        self._lib.orc_final_comm(self._h, _p32(out))
        return out

    def final_modularity(self):
        return self._lib.orc_final_modularity(self._h, _p32(self._src), _p32(self._dst),
                                              self._w.ctypes.data_as(ct.c_void_p), self._wtype)


def run(n, src, dst, w, ids=None, gamma=1.0, min_gain=1e-7, max_passes=0, trace=False,
        threads=1) -> OracleRun:
    """Run the oracle's full multi-level Louvain on an edge list.

    ``src``/``dst`` hold node ids (positions when ``ids`` is None), ``w`` is an
    int32 or float32 array.  ``trace`` keeps every pass's ``next[]``.
    """
    lib = _load()
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    w, wtype = _weights(w)
    ids_arr = None if ids is None else np.ascontiguousarray(ids, np.int32)
    h = lib.orc_run_louvain(int(n), None if ids_arr is None else _p32(ids_arr), len(src),
                            _p32(src), _p32(dst), w.ctypes.data_as(ct.c_void_p), wtype,
                            float(gamma), float(min_gain), int(max_passes), int(bool(trace)),
                            int(threads))
    if not h:
        raise OracleError("allocation failed")
    return OracleRun(h, int(n), src, dst, w, wtype, ids_arr)


def modularity(n, src, dst, w, comm, ids=None, gamma=1.0) -> float:
    """FP64 Newman-Girvan modularity of ``comm`` (indexed by position)."""
    lib = _load()
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    w, wtype = _weights(w)
    comm = np.ascontiguousarray(comm, np.int32)
    ids_arr = None if ids is None else np.ascontiguousarray(ids, np.int32)
    return lib.orc_modularity_fp64(int(n), None if ids_arr is None else _p32(ids_arr), len(src),
                                   _p32(src), _p32(dst), w.ctypes.data_as(ct.c_void_p), wtype,
                                   _p32(comm), float(gamma))
