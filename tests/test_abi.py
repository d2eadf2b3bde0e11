"""The C-ABI libraries load and export every function include/*.h declares;
the error paths that need no GPU behave (no compute calls here)."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

from louvaingpu_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(%s\w+)\s*\(" % prefix, text)))


def test_engine_exports_every_declared_symbol():
    names = declared("louvain_amd.h", "lv_")
    assert len(names) >= 25
    lib = _lib.engine()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.ENGINE_SYMBOLS)


def test_host_exports_every_declared_symbol():
    names = declared("louvain_host_c.h", "lv[hg]_")
    lib = _lib.host()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.HOST_SYMBOLS)


def test_server_exports_every_declared_symbol():
    names = declared("louvain_serve.h", "lvs_")
    lib = _lib.host()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SERVE_SYMBOLS)


def test_server_argument_errors_without_device():
    lib = _lib.host()
    assert lib.lvs_destroy(None) == _lib.LV_ERR_ARG
    assert lib.lvs_create(None, None, 1, 0) == _lib.LV_ERR_ARG
    h = ct.c_void_p()
    assert lib.lvs_create(ct.byref(h), None, 0, 0) == _lib.LV_ERR_ARG
    rid = ct.c_int64()
    assert lib.lvs_submit(None, 1, 0, None, None, None, 0, 1.0, ct.byref(rid)) == _lib.LV_ERR_ARG


def test_strerror_and_defaults():
    assert "invalid input" in _lib.strerror(_lib.LV_ERR_INVALID_INPUT)
    assert _lib.strerror(12345) == "unknown status"
    cfg = _lib.LvConfig()
    _lib.engine().lv_config_default(ct.byref(cfg))
    assert cfg.gamma == 1.0 and cfg.min_gain == 1e-7 and cfg.max_passes == 0
    assert cfg.upload == 0 and cfg.fanout == 0


def test_null_context_is_rejected():
    lib = _lib.engine()
    assert lib.lv_destroy(None) == _lib.LV_ERR_ARG
    assert lib.lv_run(None, None) == _lib.LV_ERR_ARG
    assert lib.lv_create(None, None) == _lib.LV_ERR_ARG


def test_bad_config_is_rejected_before_device_use():
    lib = _lib.engine()
    cfg = _lib.LvConfig()
    lib.lv_config_default(ct.byref(cfg))
    cfg.gamma = -1.0
    h = ct.c_void_p()
    assert lib.lv_create(ct.byref(h), ct.byref(cfg)) == _lib.LV_ERR_ARG
    for field, bad in (("hub_q", 63), ("hub_q", 4097), ("hub_part_target", 10),
                       ("chunk_mul", 2000), ("upload", 3), ("upload", -1), ("fanout", 3),
                       ("fanout", -1)):
        lib.lv_config_default(ct.byref(cfg))
        assert cfg.hub_q == cfg.hub_part_target == cfg.chunk_mul == 0
        setattr(cfg, field, bad)
        assert lib.lv_create(ct.byref(h), ct.byref(cfg)) == _lib.LV_ERR_ARG, field


def test_static_run_rejects_before_device_use():
    lib = _lib.engine()
    a = np.zeros(4, np.int32)
    p = a.ctypes.data_as(ct.POINTER(ct.c_int32))
    big_e = _lib.LV_STATIC_SCRATCH_BYTES // 200 + 1
    # beyond the static array (the arrays are not read), then argument errors
    assert lib.lv_run_static(None, None, 4, None, big_e, p, p, p, 0, p, None, None) == \
        _lib.LV_ERR_TOO_LARGE
    assert lib.lv_run_static(None, None, 4, None, 2, None, p, p, 0, p, None, None) == _lib.LV_ERR_ARG
    assert lib.lv_run_static(None, None, 4, None, 2, p, p, p, 7, p, None, None) == _lib.LV_ERR_ARG
    cfg = _lib.LvConfig()
    lib.lv_config_default(ct.byref(cfg))
    cfg.flags = _lib.LV_FLAG_TIMING
    assert lib.lv_run_static(ct.byref(cfg), None, 4, None, 2, p, p, p, 0, p, None, None) == \
        _lib.LV_ERR_ARG
    assert "static scratch" in _lib.strerror(_lib.LV_ERR_TOO_LARGE)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "_engine", None)
    monkeypatch.setattr(_lib, "ENGINE_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(_lib.NativeLibraryMissing):
        _lib.engine()
