"""The C ABI library loads and exports every symbol include/tcw_hip.h declares
(no compute without a GPU)."""

import ctypes
import os
import re

import pytest

from pyfstat_amd import _hiplib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(REPO, "include", "tcw_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tcw_[a-z_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = _hiplib.load()
    names = _declared()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_hiplib.EXPORTED_SYMBOLS)


def test_version_and_shape_without_device():
    lib = _hiplib.load()
    assert b"gfx950" in lib.tcw_version()
    w = _hiplib.window_struct((1, 1000, 10 * 1800, 1800, 3600, 20 * 1800, 1800))
    a, b = ctypes.c_int64(), ctypes.c_int64()
    assert lib.tcw_map_shape(100, 1800, 1000, ctypes.byref(w), ctypes.byref(a),
                             ctypes.byref(b)) == 0
    assert (a.value, b.value) == (11, 21)
    # TRANSIENT_NONE: one cell
    w0 = _hiplib.window_struct((0, 0, 0, 0, 0, 0, 0))
    assert lib.tcw_map_shape(100, 1800, 1000, ctypes.byref(w0), ctypes.byref(a),
                             ctypes.byref(b)) == 0
    assert (a.value, b.value) == (1, 1)
    bad = _hiplib.window_struct((3, 0, 0, 1, 0, 0, 1))
    rc = lib.tcw_map_shape(100, 1800, 1000, ctypes.byref(bad), ctypes.byref(a), ctypes.byref(b))
    assert rc == _hiplib.TCW_ERR_INVALID
    assert b"window-type" in lib.tcw_last_error()
    with pytest.raises(ValueError):
        _hiplib.check(rc)
    zero = _hiplib.window_struct((1, 0, 10, 0, 0, 0, 1))
    rc = lib.tcw_map_shape(100, 1800, 1000, ctypes.byref(zero), ctypes.byref(a), ctypes.byref(b))
    assert rc == _hiplib.TCW_ERR_INVALID


def test_create_rejects_bad_device():
    lib = _hiplib.load()
    h = ctypes.c_void_p()
    n = lib.tcw_device_count()
    rc = lib.tcw_create(n + 3, ctypes.byref(h))
    assert rc == _hiplib.TCW_ERR_INVALID
    assert b"exceeds" in lib.tcw_last_error()


def test_stats_struct_layout():
    # tcw_stats as declared (the library static_asserts the same sizes):
    # 4+4 + 6*8 + 8 + 2*8 + 2*8 + 4+4 = 104 bytes
    assert ctypes.sizeof(_hiplib.Stats) == 104
    assert ctypes.sizeof(_hiplib.Atoms) == 4 + 4 + 8 + 8 + 8 + 8 * 11 + 4 + 4
    # tcw_sft_set: n_sft, two pointers, 4 x 32-bit, four pointers
    assert ctypes.sizeof(_hiplib.SftSet) == 8 + 2 * 8 + 4 * 4 + 4 * 8
