"""Host mirrors in page-locked memory (vkh_host_*): uploads and downloads are
bit-exact round trips, a view of the mirror stays valid after its Array is
gone, and the block returns to the pinned pool only when the last view dies."""
import ctypes
import gc

import numpy as np
import pytest

import vulkpy_amd as vk
from vulkpy_amd import _lib

pytestmark = pytest.mark.gpu


def host_stats():
    live, cached = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.call("vkh_host_stats", ctypes.byref(live), ctypes.byref(cached))
    return live.value, cached.value


def test_pinned_round_trip(gpu):
    rng = np.random.default_rng(40)
    for shape in ((1 << 18,), (1000, 777), (3, 1 << 19)):   # 1 MiB, ~3 MiB, 6 MiB
        x = rng.standard_normal(shape).astype(np.float32)
        live0, _ = host_stats()
        a = vk.Array(gpu, data=x)
        assert host_stats()[0] >= live0 + x.nbytes        # the mirror is pinned
        np.testing.assert_array_equal(np.asarray(a).view(np.uint32), x.view(np.uint32))
        b = a * 2.0
        np.testing.assert_array_equal(np.asarray(b), x * np.float32(2))
        a[0] = 5.0                                       # write through the mirror
        assert np.asarray(a).reshape(-1)[0] == 5.0
        del a, b
        gc.collect()


def test_chunked_upload(gpu):
    """Mirrors of 32 MiB and more upload in 16 MiB chunks, each DMA overlapping
    the host copy of the next chunk (ragged last chunk, and a non-contiguous
    source, which takes the single-copy path)."""
    rng = np.random.default_rng(41)
    x = rng.standard_normal(9_000_001).astype(np.float32)
    a = vk.Array(gpu, data=x)
    np.testing.assert_array_equal(np.asarray(a).view(np.uint32), x.view(np.uint32))
    np.testing.assert_array_equal(np.asarray(a + 1.0), x + np.float32(1))
    t = rng.standard_normal((2000, 4500)).astype(np.float32).T   # Fortran-ordered view
    b = vk.Array(gpu, data=t)
    np.testing.assert_array_equal(np.asarray(b), t)
    u = vk.U32Array(gpu, data=np.arange(9_000_000, dtype=np.uint32))
    np.testing.assert_array_equal(np.asarray(u), np.arange(9_000_000, dtype=np.uint32))


def test_view_outlives_array(gpu):
    x = np.arange(1 << 20, dtype=np.float32)
    a = vk.Array(gpu, data=x)
    v = a.array[10:20]                                   # view into the pinned mirror
    live_with_view = host_stats()[0]
    del a
    gc.collect()
    assert host_stats()[0] == live_with_view             # still held by the view
    np.testing.assert_array_equal(v, x[10:20])
    del v
    gc.collect()
    assert host_stats()[0] <= live_with_view - x.nbytes  # returned to the pool


def test_u32_and_small_arrays(gpu):
    idx = np.arange(1 << 19, dtype=np.uint32)[::-1].copy()
    u = vk.U32Array(gpu, data=idx)
    np.testing.assert_array_equal(np.asarray(u), idx)
    small = np.float32([1.5, -2.0, 3.25])                # below the pinning threshold
    live0 = host_stats()[0]
    s = vk.Array(gpu, data=small)
    assert host_stats()[0] == live0
    np.testing.assert_array_equal(np.asarray(s), small)


def test_host_pool_errors():
    with pytest.raises(ValueError):
        _lib.call("vkh_host_free", ctypes.c_void_p(0x1000))
    _lib.call("vkh_host_empty_cache")
    assert host_stats()[1] == 0
