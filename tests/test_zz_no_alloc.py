"""lv_run_static (no device allocation, one static scratch array per GPU)
against the allocating entry points on the same inputs, through the C ABI:
two calls in a row, the second on a smaller graph, each equal bit for bit to
lv_load_edges + lv_run + lv_final_assignment / lv_final_modularity of a
context of its own."""
import ctypes as ct

import numpy as np
import pytest

from conftest import gpu_case
from louvaingpu_amd import Engine, _lib, generators
from louvaingpu_amd.engine import make_config, weights_array

pytestmark = pytest.mark.gpu


def _p32(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_int32))


def run_static(n, src, dst, w, gamma=1.0, stream=None, ids=None):
    lib = _lib.engine()
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    w, wtype = weights_array(w)
    cfg = make_config(gamma=gamma)
    out = np.full(n, -1, np.int32)
    levels = ct.c_int32(0)
    q = ct.c_double(0.0)
    ida = None if ids is None else np.ascontiguousarray(ids, np.int32)
    rc = lib.lv_run_static(ct.byref(cfg), stream, n, None if ida is None else _p32(ida), len(src),
                           _p32(src), _p32(dst), w.ctypes.data_as(ct.c_void_p), wtype, _p32(out),
                           ct.byref(levels), ct.byref(q))
    return rc, out, levels.value, q.value


def run_ctx(n, src, dst, w, gamma=1.0, ids=None):
    with Engine(gamma=gamma) as eng:
        eng.load_edges(n, src, dst, w, ids=ids)
        levels = eng.run()
        return eng.final_assignment(), levels, eng.final_modularity()


def check_same(g, gamma=1.0, stream=None, ids=None):
    rc, fin, levels, q = run_static(*g, gamma=gamma, stream=stream, ids=ids)
    assert rc == _lib.LV_OK, _lib.strerror(rc)
    efin, elevels, eq = run_ctx(*g, gamma=gamma, ids=ids)
    assert levels == elevels
    assert np.array_equal(fin, efin)
    assert q == eq  # same exact aggregates, same FP64 expression


def test_static_run_twice_second_smaller():
    big = generators.rmat(15, 16, 0.57, 0.19, 0.19, 3, "float")
    small = generators.sbm(2000, 100, 0.08, 0.0016, 7)
    assert len(big[1]) * 200 + big[0] * 320 < _lib.LV_STATIC_SCRATCH_BYTES
    gpu_case("lv_run_static rmat15 float n=%d e=%d" % (big[0], len(big[1])))
    check_same(big)
    gpu_case("lv_run_static sbm2000 n=%d e=%d gamma 1.5" % (small[0], len(small[1])))
    check_same(small, gamma=1.5)


def test_static_run_on_a_caller_stream_with_ids():
    hip = ct.CDLL("libamdhip64.so")
    s = ct.c_void_p()
    assert hip.hipStreamCreate(ct.byref(s)) == 0
    try:
        rng = np.random.Generator(np.random.PCG64(41))
        n = 3000
        ids = (rng.permutation(n) * 3 + 11).astype(np.int32)
        e = 20000
        a = rng.integers(0, n, e)
        b = rng.integers(0, n, e)
        w = rng.integers(0, 9, e).astype(np.int32)
        g = (n, ids[a], ids[b], w)
        gpu_case("lv_run_static on a created stream, arbitrary ids n=%d e=%d" % (n, e))
        check_same(g, stream=s, ids=ids)
        gpu_case("lv_run_static again, 500-vertex prefix graph")
        keep = (a < 500) & (b < 500)
        ids2 = ids[:500]
        g2 = (500, ids[a[keep]], ids[b[keep]], w[keep])
        check_same(g2, stream=s, ids=ids2)
    finally:
        assert hip.hipStreamDestroy(s) == 0


def test_static_run_rejects_oversized_and_bad_input():
    gpu_case("lv_run_static error paths")
    n = 100
    src = np.array([0, 1], np.int32)
    dst = np.array([1, 200], np.int32)  # unknown vertex
    rc, _, _, _ = run_static(n, src, dst, np.ones(2, np.int32))
    assert rc == _lib.LV_ERR_INVALID_INPUT
    # beyond the array: rejected before the arrays are read
    big_e = _lib.LV_STATIC_SCRATCH_BYTES // 200 + 1
    lib = _lib.engine()
    cfg = make_config()
    out = np.zeros(n, np.int32)
    rc = lib.lv_run_static(ct.byref(cfg), None, n, None, big_e, _p32(src), _p32(dst),
                           src.ctypes.data_as(ct.c_void_p), _lib.LV_WEIGHT_INT32, _p32(out),
                           None, None)
    assert rc == _lib.LV_ERR_TOO_LARGE
    # a valid call afterwards still matches
    g = generators.sbm(1000, 100, 0.1, 0.002, 3)
    check_same(g)
