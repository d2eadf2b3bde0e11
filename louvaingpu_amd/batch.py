"""Graph sets: many independent graphs, each at several resolutions, in one
engine run (``lv_batch_*`` in ``include/louvain_amd.h``).

The reference runs one graph per ``generate_dendogram`` call
(``LouvainCUDA/LouvainCPP.cpp:209-553``).  ``GraphSet`` packs a list of graphs
times a list of resolutions into the components of one union graph on the
device, so every phase is one launch per level and pass for the whole set;
each component's result equals a single-graph run of that graph at that
resolution.  Component ``c`` is graph ``c // len(gammas)`` at resolution
``gammas[c % len(gammas)]``.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from ._lib import LvLevelInfo, check
from .engine import Engine, _p32, _p64, weights_array


class GraphSet:
    """A loaded set of graphs x resolutions on one engine context."""

    def __init__(self, engine: Engine | None = None, **config):
        self._own = engine is None
        self.engine = engine if engine is not None else Engine(**config)
        self._lib = self.engine._lib
        self.ngraphs = 0
        self.nres = 0
        self.n = None

    def close(self):
        if self._own and self.engine is not None:
            self.engine.close()
        self.engine = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def _h(self):
        return self.engine._h

    def load(self, graphs, gammas):
        """graphs: list of (n, src, dst, w) with ids 0..n-1 and one weight dtype
        (int32 or float32) for all; gammas: the resolutions every graph runs at."""
        if not graphs:
            raise ValueError("empty graph list")
        ns = np.array([int(g[0]) for g in graphs], np.int32)
        es = np.array([len(g[1]) for g in graphs], np.int64)
        srcs = [np.ascontiguousarray(g[1], np.int32) for g in graphs]
        dsts = [np.ascontiguousarray(g[2], np.int32) for g in graphs]
        ws = [np.ascontiguousarray(g[3]) for g in graphs]
        if any(not (len(a) == len(b) == len(c)) for a, b, c in zip(srcs, dsts, ws)):
            raise ValueError("src, dst and w of a graph must have the same length")
        dt = ws[0].dtype
        if any(w.dtype != dt for w in ws):
            raise TypeError("all graphs must share one weight dtype")
        _, wtype = weights_array(ws[0])
        vp = ct.c_void_p * len(graphs)
        ps = vp(*[a.ctypes.data for a in srcs])
        pd = vp(*[a.ctypes.data for a in dsts])
        pw = vp(*[a.ctypes.data for a in ws])
        gam = np.ascontiguousarray(gammas, np.float64)
        check(self._lib.lv_batch_load_graphs(self._h, len(graphs), _p32(ns), _p64(es), ps, pd, pw,
                                             wtype, len(gam),
                                             gam.ctypes.data_as(ct.POINTER(ct.c_double))),
              "lv_batch_load_graphs")
        self.ngraphs, self.nres, self.n = len(graphs), len(gam), ns
        return self

    def run(self) -> float:
        """Runs every component's level loop; returns the host wall time in ms."""
        check(self._lib.lv_batch_run(self._h), "lv_batch_run")
        return self.engine.last_run_ms()

    @property
    def ncomp(self) -> int:
        return self.ngraphs * self.nres

    def comp_n(self, comp: int) -> int:
        return int(self.n[comp // self.nres])

    def num_levels(self, comp: int) -> int:
        v = ct.c_int32()
        check(self._lib.lv_batch_num_levels(self._h, int(comp), ct.byref(v)),
              "lv_batch_num_levels")
        return v.value

    def level_info(self, comp: int, level: int) -> LvLevelInfo:
        out = LvLevelInfo()
        check(self._lib.lv_batch_level_info(self._h, int(comp), int(level), ct.byref(out)),
              "lv_batch_level_info")
        return out

    def level_assignment(self, comp: int, level: int) -> np.ndarray:
        nv = self.level_info(comp, level).n
        out = np.empty(nv, np.int32)
        check(self._lib.lv_batch_level_assignment(self._h, int(comp), int(level), _p32(out)),
              "lv_batch_level_assignment")
        return out

    def final(self, comp: int) -> np.ndarray:
        out = np.empty(self.comp_n(comp), np.int32)
        check(self._lib.lv_batch_final_assignment(self._h, int(comp), _p32(out)),
              "lv_batch_final_assignment")
        return out

    def final_all(self) -> np.ndarray:
        """Final assignments of all components, concatenated in component order."""
        total = int(self.n.astype(np.int64).sum()) * self.nres
        out = np.empty(total, np.int32)
        check(self._lib.lv_batch_final_assignment(self._h, -1, _p32(out)),
              "lv_batch_final_assignment")
        return out
