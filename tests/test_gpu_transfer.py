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


def test_mirror_survives_inplace_then_delete(gpu):
    """An upload from a pinned mirror still queued behind a long kernel, an
    in-place op on the array, then the array deleted: its mirror goes back to
    the pool only after the upload ran, so a new array of the same size class
    (which takes that block and writes other bytes into it) cannot change
    what the first array's consumer computes."""
    n = 1 << 22                                            # 16 MiB: pinned, one chunk
    rng = np.random.default_rng(42)
    first = rng.standard_normal(n).astype(np.float32)
    second = rng.standard_normal(n).astype(np.float32)
    big = vk.Array(gpu, shape=(4096, 4096))
    big[:] = 1.0
    gpu.wait()
    slow = big @ big                                       # long kernel ahead in the stream
    x = vk.Array(gpu, data=first)                          # upload queued behind it
    x *= 2.0                                               # in-place: new ticket on x
    acc = x.sum()                                          # consumer of x's bytes
    del x
    gc.collect()
    y = vk.Array(gpu, data=second)                         # same size class: may reuse the block
    got = float(np.asarray(acc))
    want = float((first.astype(np.float64) * 2).sum())
    assert abs(got - want) <= 1e-6 * np.abs(first).sum() * 2, (got, want)
    np.testing.assert_array_equal(np.asarray(y), second)
    del slow, big


def test_upload_download_refused_during_capture(gpu):
    """Copies to or from host memory are not graph nodes: a captured upload
    would read its host buffer again at every replay."""
    h = gpu._handle
    dev = vk.Array(gpu, shape=(16,))
    host = np.zeros(16, dtype=np.float32)
    gpu.wait()
    _lib.call("vkh_graph_begin", h)
    try:
        assert _lib.lib.vkh_upload(h, ctypes.c_void_p(dev._ptr), host.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.c_size_t(64)) == _lib.VKH_INVALID
        assert "graph capture" in _lib.last_error()
        assert _lib.lib.vkh_download(h, host.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(dev._ptr),
                                     ctypes.c_size_t(64)) == _lib.VKH_INVALID
        # nor are event records (a job wait): the node would outlive its event
        ev = ctypes.c_void_p()
        assert _lib.lib.vkh_event_record(h, ctypes.byref(ev)) == _lib.VKH_INVALID
        assert _lib.lib.vkh_timer_record(h, ctypes.byref(ev)) == _lib.VKH_INVALID
        assert ev.value is None
        _lib.call("vkh_fill_f32", h, dev._ptr, 1.0, 16)    # the capture itself goes on
    finally:
        ex = ctypes.c_void_p()
        _lib.call("vkh_graph_end", h, ctypes.byref(ex))
    dev[:] = 3.0
    np.testing.assert_array_equal(np.asarray(dev), np.full(16, 3.0, np.float32))
    _lib.call("vkh_graph_launch", h, ex.value)
    np.testing.assert_array_equal(_download(gpu, dev),
                                  np.ones(16, np.float32))
    _lib.call("vkh_graph_destroy", ex.value)


def _download(gpu, arr):
    out = np.empty(arr.size, dtype=np.float32)
    _lib.call("vkh_download", gpu._handle, out.ctypes.data, arr._ptr, out.nbytes)
    return out
