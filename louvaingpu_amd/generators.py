"""Deterministic synthetic graphs for the benchmark configurations.

Thin wrappers over the C++ generators in ``csrc/host/generators.cpp``
(SplitMix64-seeded xoshiro256** streams; SBM by geometric skipping, R-MAT with
duplicate merging).  Each returns ``(n, src, dst, w)`` with canonical
(u <= v), lexicographically sorted edges.
"""
from __future__ import annotations

import ctypes as ct
import math

import numpy as np

from . import _lib


def _take(g):
    lib = _lib.host()
    if not g:
        raise ValueError("generator rejected its arguments")
    try:
        n = lib.lvg_num_nodes(g)
        e = lib.lvg_num_edges(g)
        src = np.empty(e, np.int32)
        dst = np.empty(e, np.int32)
        w = np.empty(e, np.float32 if lib.lvg_wtype(g) == 1 else np.int32)
        if e:
            lib.lvg_copy(g, src.ctypes.data_as(ct.POINTER(ct.c_int32)),
                         dst.ctypes.data_as(ct.POINTER(ct.c_int32)), w.ctypes.data_as(ct.c_void_p))
        return n, src, dst, w
    finally:
        lib.lvg_free(g)


def sbm(n, block_size, p_in, p_out, seed):
    """Planted-partition SBM: block(v) = v // block_size, unit int32 weights."""
    return _take(_lib.host().lvg_sbm(int(n), int(block_size), float(p_in), float(p_out),
                                     int(seed) & 0xFFFFFFFFFFFFFFFF))


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0, weights="int"):
    """R-MAT: int32 weights count merged draws; float32 weights sum U[1,10] draws."""
    wtype = {"int": 0, "float": 1}[weights]
    return _take(_lib.host().lvg_rmat(int(scale), int(edge_factor), float(a), float(b),
                                      float(c), int(seed) & 0xFFFFFFFFFFFFFFFF, wtype))


# The configurations the benchmark and the parity tests use.
def sbm_10k(seed=0):
    """10,000 vertices in 100 equal communities, p_in 0.05, p_out 0.0005."""
    return sbm(10000, 100, 0.05, 0.0005, seed)


def rmat22(seed=1):
    """R-MAT scale 22, edge factor 16, a/b/c/d = .57/.19/.19/.05, int32 weights."""
    return rmat(22, 16, 0.57, 0.19, 0.19, seed, "int")


def rmat26_float(seed=2):
    """R-MAT scale 26, edge factor 16, float32 weights U[1,10] summed over duplicates."""
    return rmat(26, 16, 0.57, 0.19, 0.19, seed, "float")


def batch_params(count=4096, seed=3):
    """Per-graph (n, p_in, seed) of the batched workload: n log-uniform in
    [100, 50000], p_in ~ U[0.02, 0.1], p_out = p_in / 50, blocks of 100."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = np.floor(np.exp(rng.uniform(math.log(100), math.log(50000), count))).astype(np.int64)
    pin = rng.uniform(0.02, 0.1, count)
    seeds = rng.integers(0, 2**62, count)
    return [(int(ns[i]), float(pin[i]), int(seeds[i])) for i in range(count)]


def resolutions(k=16):
    """gamma_j = 0.25 * 16**(j/15), j = 0..15 (0.25 .. 4.0 geometric)."""
    return [0.25 * 16 ** (j / (k - 1)) for j in range(k)]
