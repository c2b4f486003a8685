"""vkh_reduce_all_noalloc: the full reduction with its FP64 block partials in
the library's file-scope device array instead of a workspace.  Called through
the C-ABI, it gives vkh_reduce_all's results bit for bit, on two calls in a
row (the second on a smaller input, so it runs over partials the first call
left), for every op and for the mean, and the context's allocator holds the
same live and cached bytes before and after the calls."""
import ctypes

import numpy as np
import pytest

import vulkpy_amd as vk
from vulkpy_amd import _lib
from vulkpy_amd._lib import REDOP

pytestmark = pytest.mark.gpu


def download(gpu, arr):
    out = np.empty(arr.size, dtype=np.float32)
    _lib.call("vkh_download", gpu._handle, out.ctypes.data, arr._ptr, out.nbytes)
    return out


def test_noalloc_equals_reduce_all(gpu):
    rng = np.random.default_rng(90)
    h = gpu._handle
    cases = []
    for n in ((1 << 24) + 3, 1000):                       # second call: smaller input
        x = rng.standard_normal(n).astype(np.float32)
        p = rng.uniform(0.999, 1.001, n).astype(np.float32)
        cases.append((n, vk.Array(gpu, data=x), vk.Array(gpu, data=p)))
    runs = [(op, mean) for op in ("sum", "maximum", "minimum", "prod") for mean in (0, 1)
            if mean == 0 or op == "sum"]
    outs = {(n, op, mean, which): vk.Array(gpu, shape=(1,))
            for n, _, _ in cases for op, mean in runs for which in (0, 1)}
    gpu.wait()
    before = gpu.memory_stats()
    for n, X, P in cases:
        for op, mean in runs:
            src = P if op == "prod" else X
            _lib.call("vkh_reduce_all_noalloc", h, REDOP[op], src._ptr, n, outs[(n, op, mean, 1)]._ptr, mean)
    gpu.wait()
    assert gpu.memory_stats() == before                   # nothing allocated, nothing cached
    for n, X, P in cases:
        for op, mean in runs:
            src = P if op == "prod" else X
            _lib.call("vkh_reduce_all", h, REDOP[op], src._ptr, n, outs[(n, op, mean, 0)]._ptr, mean)
    for n, X, P in cases:
        for op, mean in runs:
            want = download(gpu, outs[(n, op, mean, 0)])
            got = download(gpu, outs[(n, op, mean, 1)])
            assert got.view(np.uint32)[0] == want.view(np.uint32)[0], (n, op, mean, got, want)
    # and against the values themselves
    x = download(gpu, cases[1][1])
    assert download(gpu, outs[(1000, "maximum", 0, 1)])[0] == x.max()
    s = float(download(gpu, outs[(1000, "sum", 0, 1)])[0])
    assert abs(s - x.astype(np.float64).sum()) <= np.spacing(np.float32(abs(s)))


def test_noalloc_repeated_small_then_empty(gpu):
    """A call over fewer blocks after one over many, then a zero-size sum."""
    h = gpu._handle
    x = np.arange(1, 70001, dtype=np.float32)
    X = vk.Array(gpu, data=x)
    big = vk.Array(gpu, data=np.ones(1 << 22, dtype=np.float32))
    out = vk.Array(gpu, shape=(1,))
    _lib.call("vkh_reduce_all_noalloc", h, REDOP["sum"], big._ptr, 1 << 22, out._ptr, 0)
    assert download(gpu, out)[0] == float(1 << 22)
    _lib.call("vkh_reduce_all_noalloc", h, REDOP["sum"], X._ptr, 70000, out._ptr, 0)
    assert download(gpu, out)[0] == np.float32(70000 * 70001 / 2)
    _lib.call("vkh_reduce_all_noalloc", h, REDOP["sum"], X._ptr, 0, out._ptr, 0)
    assert download(gpu, out)[0] == 0.0
    assert _lib.lib.vkh_reduce_all_noalloc(h, REDOP["maximum"], ctypes.c_void_p(X._ptr), 0,
                                           ctypes.c_void_p(out._ptr), 0) == _lib.VKH_INVALID
