"""Serving: continuous batching of Louvain requests (``include/louvain_serve.h``).

``Server`` wraps the C++ request loop of ``liblouvain_host.so``: ``submit``
queues a graph without copying it (the server keeps references to the arrays
until ``wait`` returns, since the worker reads them when it uploads the set),
worker threads (one engine context and HIP stream each) admit waiting requests
by HBM footprint into graph sets and run them while later requests queue, and
``wait`` returns a request's final partition with its timings, or raises
``LouvainError`` with that request's own status.  A graph too large for the
budget is refused by ``submit`` (``LV_ERR_NOMEM``).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from ._lib import LvsResult, LvsStats, check
from .engine import _p32, make_config, weights_array


class Server:
    def __init__(self, workers=2, budget_bytes=0, **config):
        self._lib = _lib.host()
        cfg = make_config(**config)
        h = ct.c_void_p()
        check(self._lib.lvs_create(ct.byref(h), ct.byref(cfg), int(workers), int(budget_bytes)),
              "lvs_create")
        self._h = h
        self._n = {}
        self._keep = {}  # the arrays a queued request borrows

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lvs_destroy(self._h)  # finishes queued requests first
        self._h = None
        self._keep = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, n, src, dst, w, gamma=1.0) -> int:
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        w, wtype = weights_array(w)
        rid = ct.c_int64()
        check(self._lib.lvs_submit(self._h, int(n), len(src), _p32(src), _p32(dst),
                                   w.ctypes.data_as(ct.c_void_p), wtype, float(gamma),
                                   ct.byref(rid)), "lvs_submit")
        self._n[rid.value] = int(n)
        self._keep[rid.value] = (src, dst, w)
        return rid.value

    def wait(self, rid: int):
        """(final partition, LvsResult) of a request; forgets it afterwards."""
        n = self._n.pop(rid)
        comm = np.empty(n, np.int32)
        res = LvsResult()
        try:
            check(self._lib.lvs_wait(self._h, int(rid), _p32(comm), ct.byref(res)), "lvs_wait")
        finally:
            self._keep.pop(rid, None)
        return comm, res

    def now_ms(self) -> float:
        return self._lib.lvs_now_ms(self._h)

    def stats(self) -> LvsStats:
        st = LvsStats()
        check(self._lib.lvs_get_stats(self._h, ct.byref(st)), "lvs_get_stats")
        return st
