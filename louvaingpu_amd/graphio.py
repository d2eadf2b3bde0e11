"""Graph files over the host library's C ABI (``include/louvain_host_c.h``).

* ``load_text(nodes_path, edges_path)``: the reference CLI's ``nodes.txt`` /
  ``edges.txt`` (``LouvainCUDA/LouvainCPP.cpp:660-737``): a count, then node
  ids, or ``source target weight`` triples; float weights.  A node id that ends
  the file without trailing whitespace is not read, as in the reference.
* ``save_text``: writes that format (with trailing newlines).
* ``load_binary`` / ``save_binary``: the little-endian binary edge list
  (32-byte header, then ids, src, dst, w as 4-byte arrays) for graphs too
  large for text; the arrays are read straight into numpy buffers.

Every loader returns ``(n, src, dst, w, ids)`` with ``ids`` None when the file
stores no node ids (then the endpoints are positions 0..n-1).
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from . import _lib
from ._lib import check
from .engine import _p32, weights_array


def _path(p):
    return os.fsencode(os.fspath(p))


def load_text(nodes_path, edges_path):
    lib = _lib.host()
    h = ct.c_void_p()
    check(lib.lvh_load_text(_path(nodes_path), _path(edges_path), ct.byref(h)), "lvh_load_text")
    try:
        n = lib.lvh_file_num_nodes(h)
        e = lib.lvh_file_num_edges(h)
        ids = np.empty(n, np.int32)
        src = np.empty(e, np.int32)
        dst = np.empty(e, np.int32)
        w = np.empty(e, np.float32)
        check(lib.lvh_file_copy(h, _p32(ids), _p32(src), _p32(dst),
                                w.ctypes.data_as(ct.POINTER(ct.c_float))), "lvh_file_copy")
        return n, src, dst, w, ids
    finally:
        lib.lvh_file_free(h)


def save_text(nodes_path, edges_path, ids, src, dst, w):
    """nodes.txt: the count, then one id per line; edges.txt: the count, then
    one `source target weight` line per edge (ids, not positions)."""
    ids = np.asarray(ids)
    with open(nodes_path, "w") as f:
        f.write("%d\n" % len(ids))
        f.write("".join("%d\n" % v for v in ids.tolist()))
    with open(edges_path, "w") as f:
        f.write("%d\n" % len(src))
        ws = [repr(float(x)) for x in np.asarray(w, np.float32).tolist()]
        f.write("".join("%d %d %s\n" % (a, b, c)
                        for a, b, c in zip(np.asarray(src).tolist(), np.asarray(dst).tolist(), ws)))


def load_binary(path):
    lib = _lib.host()
    n, e = ct.c_int32(), ct.c_int64()
    wtype, has_ids = ct.c_int(), ct.c_int()
    p = _path(path)
    check(lib.lvh_binary_header(p, ct.byref(n), ct.byref(e), ct.byref(wtype), ct.byref(has_ids)),
          "lvh_binary_header")
    ids = np.empty(n.value, np.int32) if has_ids.value else None
    src = np.empty(e.value, np.int32)
    dst = np.empty(e.value, np.int32)
    w = np.empty(e.value, np.float32 if wtype.value == _lib.LV_WEIGHT_FLOAT32 else np.int32)
    check(lib.lvh_binary_read(p, None if ids is None else _p32(ids), _p32(src), _p32(dst),
                              w.ctypes.data_as(ct.c_void_p)), "lvh_binary_read")
    return n.value, src, dst, w, ids


def save_binary(path, n, src, dst, w, ids=None):
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    w, wtype = weights_array(w)
    if not (len(src) == len(dst) == len(w)):
        raise ValueError("src, dst and w must have the same length")
    ida = None if ids is None else np.ascontiguousarray(ids, np.int32)
    if ida is not None and len(ida) != n:
        raise ValueError("ids must have n entries")
    check(_lib.host().lvh_binary_write(_path(path), int(n), None if ida is None else _p32(ida),
                                       len(src), _p32(src), _p32(dst),
                                       w.ctypes.data_as(ct.c_void_p), wtype), "lvh_binary_write")
