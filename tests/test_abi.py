"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/vulkpy_hip.h, and the Python binding declares each one."""
import ctypes
import os
import re

import pytest

from vulkpy_amd import _lib


def _header_functions():
    text = open(_lib.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vkh_\w+)\s*\(", text)))


def test_header_parses():
    names = _header_functions()
    assert "vkh_gpu_create" in names and "vkh_gemm" in names
    assert len(names) > 50


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _header_functions() if not hasattr(lib, n)]
    assert missing == []


def test_binding_declares_every_function():
    declared = set(_lib.exported_symbols())
    assert set(_header_functions()) == declared


def test_abi_version():
    v = ctypes.c_int()
    assert _lib.lib.vkh_abi_version(ctypes.byref(v)) == 0
    assert v.value == 1


def test_error_paths_without_device():
    # no compute call is made: only argument validation and device discovery
    n = ctypes.c_int(-1)
    assert _lib.lib.vkh_device_count(ctypes.byref(n)) == 0
    if n.value == 0:
        out = ctypes.c_void_p()
        st = _lib.lib.vkh_gpu_create(0, 0.0, ctypes.byref(out))
        assert st == _lib.VKH_NO_DEVICE
        assert "no HIP device" in _lib.last_error()
        import vulkpy_amd as vk
        with pytest.raises(RuntimeError):
            vk.GPU()
    # null handles are rejected, not dereferenced
    assert _lib.lib.vkh_gpu_wait(None) == _lib.VKH_INVALID
    assert _lib.lib.vkh_event_wait(None, 0) == _lib.VKH_INVALID
    assert _lib.lib.vkh_binary(None, 0, None, None, None, 0) == _lib.VKH_INVALID
    with pytest.raises(ValueError):
        _lib.check(_lib.VKH_INVALID, "x")
    with pytest.raises(RuntimeError):
        _lib.check(_lib.VKH_TIMEOUT, "x")
