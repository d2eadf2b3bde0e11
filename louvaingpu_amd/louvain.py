"""The reference's Python-facing API and dendrogram helpers.

``start(nodeList, edgeList, part_init)`` mirrors the Boost.Python entry point
(``LouvainCUDA/LouvainCPP.cpp:808-941``): node dicts with ``"id"``, edge dicts
with ``"source"``, ``"target"``, ``"weight"``; it returns
``{"size": communities, node_id: community, ...}`` from ``getPartitions`` and
writes ``"cluster"``/``"my_community"`` and the circle-layout position
``"x"``/``"y"`` of the reference's visualisation post-processing
(``calculateClusterSize`` + ``layoutProcess``, ``LouvainCUDA/LouvainCPP.cpp:555-651``)
into each node dict.  ``part_init`` is accepted and ignored, as in the reference.

``generate_dendogram`` runs the reference's level loop through the C++ host
layer (``generate_dendogram`` in ``csrc/host/louvain.cpp``), which calls the
device engine phase by phase.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from ._lib import check
from .engine import _p32, make_config


class Dendrogram:
    """Levels of (node, community) arrays in the reference's Partition layout."""

    def __init__(self, handle):
        self._lib = _lib.host()
        self._h = handle
        self.levels = []
        for l in range(self._lib.lvh_num_levels(handle)):
            cnt = self._lib.lvh_level_count(handle, l)
            node = np.empty(cnt, np.int32)
            comm = np.empty(cnt, np.int32)
            self._lib.lvh_level_partition(handle, l, node.ctypes.data_as(ct.POINTER(ct.c_int32)),
                                          comm.ctypes.data_as(ct.POINTER(ct.c_int32)))
            self.levels.append((node, comm))

    def __len__(self):
        return len(self.levels)

    def _two(self, fn, level):
        n = len(self.levels[0][0])
        node = np.empty(n, np.int32)
        comm = np.empty(n, np.int32)
        _lib.check(fn(self._h, int(level), node.ctypes.data_as(ct.POINTER(ct.c_int32)),
                      comm.ctypes.data_as(ct.POINTER(ct.c_int32))))
        return node, comm

    def partition_at_level(self, level):
        """partitionAtLevel (LouvainCUDA/LouvainCPP.cpp:82-106)."""
        return self._two(self._lib.lvh_partition_at_level, level)

    def get_partitions(self, level=0):
        """getPartitions (LouvainCUDA/LouvainCPP.cpp:108-114): level `level or len-1`
        evaluated as a boolean, i.e. level 1 whenever there are two levels."""
        return self._two(self._lib.lvh_get_partitions, level)

    def final_partition(self):
        return self.partition_at_level(len(self.levels) - 1)

    def close(self):
        if self._h:
            self._lib.lvh_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def count_partitions(comm) -> int:
    """countPartitions (LouvainCUDA/LouvainCPP.cpp:116-127)."""
    return int(np.max(comm)) + 1 if len(comm) else 1


def generate_dendogram(n, src, dst, w, ids=None, **config) -> Dendrogram:
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    wf = np.ascontiguousarray(w, np.float32)
    ids_arr = None if ids is None else np.ascontiguousarray(ids, np.int32)
    cfg = make_config(**config)
    h = ct.c_void_p()
    lib = _lib.host()
    _lib.check(lib.lvh_generate_dendrogram(
        int(n), None if ids_arr is None else ids_arr.ctypes.data_as(ct.POINTER(ct.c_int32)),
        len(src), src.ctypes.data_as(ct.POINTER(ct.c_int32)),
        dst.ctypes.data_as(ct.POINTER(ct.c_int32)), wf.ctypes.data_as(ct.POINTER(ct.c_float)),
        ct.byref(cfg), ct.byref(h)), "lvh_generate_dendrogram")
    return Dendrogram(h)


def cluster_layout(ids, node, comm, nclusters):
    """calculateClusterSize + layoutProcess (LouvainCUDA/LouvainCPP.cpp:555-651) on
    the host library: (cluster_size, cluster, index_in_cluster, x, y) for the
    nodes ``ids`` under the Partition (node, comm)."""
    lib = _lib.host()
    ids = np.ascontiguousarray(ids, np.int32)
    node = np.ascontiguousarray(node, np.int32)
    comm = np.ascontiguousarray(comm, np.int32)
    n = len(ids)
    cs = np.empty(max(int(nclusters), 1), np.int32)
    cl = np.empty(n, np.int32)
    ix = np.empty(n, np.int32)
    x = np.empty(n, np.float32)
    y = np.empty(n, np.float32)
    fp = ct.POINTER(ct.c_float)
    check(lib.lvh_cluster_layout(n, _p32(ids), len(node), _p32(node), _p32(comm), int(nclusters),
                                 _p32(cs), _p32(cl), _p32(ix), x.ctypes.data_as(fp),
                                 y.ctypes.data_as(fp)), "lvh_cluster_layout")
    return cs[:int(nclusters)], cl, ix, x, y


def start(nodeList, edgeList, part_init=None, **config):
    """Boost.Python ``start`` (LouvainCUDA/LouvainCPP.cpp:808-941)."""
    ids = np.array([int(d["id"]) for d in nodeList], np.int32)
    src = np.array([int(d["source"]) for d in edgeList], np.int32)
    dst = np.array([int(d["target"]) for d in edgeList], np.int32)
    w = np.array([float(d["weight"]) for d in edgeList], np.float32)
    d = generate_dendogram(len(ids), src, dst, w, ids=ids, **config)
    try:
        node, comm = d.get_partitions()
    finally:
        d.close()
    size = count_partitions(comm)
    result = {"size": size}
    for a, b in zip(node.tolist(), comm.tolist()):
        result[a] = b
    _, cl, _, x, y = cluster_layout(ids, node, comm, size)
    for i, nd in enumerate(nodeList):
        nd["cluster"] = int(cl[i])
        nd["my_community"] = int(cl[i])
        nd["x"] = float(x[i])
        nd["y"] = float(y[i])
    return result
