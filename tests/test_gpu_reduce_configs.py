"""Reduction kernel configurations and their autotuner (vkh_reduce_*).

Every configuration must keep the reduction's parity bar: max / min bit-exact
against the oracle, sums / products / means within k * 2^-24 of FP64 (k =
elements per output).  Configurations that cannot change the summation order
-- every ROWS configuration, and the unroll of every family -- must give
bit-identical results."""
import ctypes

import numpy as np
import pytest

import vulkpy_amd as vk
from vulkpy_amd import _lib
from vulkpy_amd.vkarray import _reduce_problems
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FAMILY_ID = {"all": 0, "rows": 1, "cols": 2}


def A(gpu, x):
    return vk.Array(gpu, data=np.asarray(x, dtype=np.float32))


def set_cfg(gpu, problem, index):
    kind, prev, K, post = problem
    _lib.call("vkh_reduce_set_config", gpu._handle, kind, prev, K, post, index)


def get_cfg(gpu, problem):
    kind, prev, K, post = problem
    idx, tuned = ctypes.c_int(), ctypes.c_int()
    _lib.call("vkh_reduce_get_config", gpu._handle, kind, prev, K, post, ctypes.byref(idx),
              ctypes.byref(tuned))
    return idx.value, tuned.value


def family(problem):
    kind, prev, K, post = problem
    f = ctypes.c_int()
    _lib.call("vkh_reduce_family", kind, prev, K, post, ctypes.byref(f))
    return f.value


def bits(x):
    return np.asarray(x, dtype=np.float32).view(np.uint32)


def test_config_tables():
    defaults = {"all": (4, 4), "rows": (4, 16), "cols": (4, 512)}
    for fam, d in defaults.items():
        cfgs = vk.reduce_configs(fam)
        assert len(cfgs) >= 4
        assert (cfgs[0]["unroll"], cfgs[0]["par"]) == d
        assert {c["unroll"] for c in cfgs} == {4, 8}
        assert len({(c["unroll"], c["par"]) for c in cfgs}) == len(cfgs)
    with pytest.raises(ValueError):
        _lib.call("vkh_reduce_configs", 7, ctypes.byref(ctypes.c_int()))


def test_families():
    assert family((0, 1, 100, 1)) == 0
    assert family((1, 10, 2048, 1)) == 1
    assert family((1, 10, 2047, 1)) == -1
    assert family((1, 1, 5, 3)) == 2
    with pytest.raises(ValueError):
        family((2, 1, 5, 3))


CASES = [
    # (shape, axis, family)
    ((1 << 22) + 5, None, "all"),
    (100003, None, "all"),
    ((37, 2048), 1, "rows"),
    ((300, 6000), 1, "rows"),
    ((5, 4099), 1, "rows"),          # row length not a multiple of 4: scalar loads
    ((4096, 64), 0, "cols"),
    ((3, 2000, 40), 1, "cols"),
    ((1000, 6), 0, "cols"),          # post not a multiple of 4: one column per thread
    ((2, 513, 4), 1, "cols"),
]


@pytest.mark.parametrize("shape,axis,fam", CASES)
def test_every_config_keeps_the_bar(gpu, shape, axis, fam):
    rng = np.random.default_rng(hash((shape, axis)) % (1 << 32))
    x = rng.uniform(0.5, 1.5, shape).astype(np.float32)
    p = rng.uniform(0.999, 1.001, shape).astype(np.float32)
    gx, gp = A(gpu, x), A(gpu, p)
    (problem,) = _reduce_problems(x.shape, axis)
    assert family(problem) == FAMILY_ID[fam]
    x64, p64 = x.astype(np.float64), p.astype(np.float64)
    sum64 = x64.sum(axis=axis)
    prod64 = p64.prod(axis=axis)
    k = x.size // np.asarray(sum64).size
    # full reductions accumulate in FP64 and round once
    tol = 2.0 ** -23 if axis is None else 4 * k * 2.0 ** -24 + 1e-7
    if axis is None:
        mx, mn = x.max(), x.min()
    else:
        mx, mn = O.reduce_axis("maximum", x, axis), O.reduce_axis("minimum", x, axis)
    results = {}
    try:
        for i, cfg in enumerate(vk.reduce_configs(fam)):
            set_cfg(gpu, problem, i)
            assert get_cfg(gpu, problem) == (i, 1)
            s = np.asarray(gx.sum(axis=axis))
            m = np.asarray(gx.mean(axis=axis))
            pr = np.asarray(gp.prod(axis=axis))
            np.testing.assert_array_equal(bits(np.asarray(gx.maximum(axis=axis)).ravel()),
                                          bits(np.ravel(mx)))
            np.testing.assert_array_equal(bits(np.asarray(gx.minimum(axis=axis)).ravel()),
                                          bits(np.ravel(mn)))
            np.testing.assert_allclose(s.ravel(), np.ravel(sum64), rtol=tol, err_msg=str(cfg))
            np.testing.assert_allclose(m.ravel(), np.ravel(sum64) / k, rtol=tol, err_msg=str(cfg))
            np.testing.assert_allclose(pr.ravel(), np.ravel(prod64), rtol=tol, err_msg=str(cfg))
            results[(cfg["unroll"], cfg["par"])] = (bits(s), bits(pr))
    finally:
        set_cfg(gpu, problem, -1)
    assert get_cfg(gpu, problem) == (0, 0)
    # configurations that cannot change the order give identical bits
    for (u, par), (s, pr) in results.items():
        if fam == "rows":
            ref = results[(4, 16)]
        else:
            ref = results[(4, par)]
        np.testing.assert_array_equal(s, ref[0], err_msg=f"unroll {u} par {par}")
        np.testing.assert_array_equal(pr, ref[1], err_msg=f"unroll {u} par {par}")


def test_rebroadcast_uses_configs(gpu):
    rng = np.random.default_rng(31)
    x = rng.uniform(0.5, 1.5, (6, 3000)).astype(np.float32)
    problem = (1, 6, 3000, 1)
    ref = O.reduce_axis("maximum", x, 1, rebroadcast=True)
    try:
        for i in range(len(vk.reduce_configs("rows"))):
            set_cfg(gpu, problem, i)
            got = A(gpu, x).maximum(axis=1, rebroadcast=True)
            np.testing.assert_array_equal(bits(got), bits(ref))
    finally:
        set_cfg(gpu, problem, -1)


def test_autotune_api(gpu):
    shape = (2048, 4096)
    rng = np.random.default_rng(32)
    x = rng.standard_normal(shape).astype(np.float32)
    g = A(gpu, x)
    tuned = []
    try:
        for axis in (0, 1, None, (0, 1)):
            res = vk.autotune_reduce(gpu, shape, axis, reps=2)
            assert len(res) == 1
            r = res[0]
            problem = r["shape"]
            fam = r["family"]
            assert fam in ("all", "rows", "cols")
            n = len(vk.reduce_configs(fam))
            assert len(r["ms"]) == n and all(t > 0 for t in r["ms"])
            assert 0 <= r["best"] < n and r["ms"][r["best"]] == min(r["ms"])
            assert r["config"] == vk.reduce_configs(fam)[r["best"]]
            assert get_cfg(gpu, problem) == (r["best"], 1)
            tuned.append(problem)
            # the tuned configuration is what later calls use
            got = np.asarray(g.sum(axis=axis))
            set_cfg(gpu, problem, r["best"])
            again = np.asarray(g.sum(axis=axis))
            np.testing.assert_array_equal(bits(got), bits(again))
        # a shape with a single kernel choice
        res = vk.autotune_reduce(gpu, (100, 50), 1)
        assert res[0]["family"] is None and res[0]["ms"] == [-1.0] and res[0]["best"] == 0
        assert get_cfg(gpu, (1, 100, 50, 1)) == (0, 0)
    finally:
        for problem in tuned:
            set_cfg(gpu, problem, -1)
    for problem in tuned:
        assert get_cfg(gpu, problem) == (0, 0)
    with pytest.raises(ValueError):
        set_cfg(gpu, (1, 1, 4096, 64), 99)
    with pytest.raises(ValueError):
        _lib.call("vkh_reduce_autotune", gpu._handle, 1, 0, 5, 5, 1, 0, None, None)


def test_autotune_same_order(gpu):
    """same_order times only the configurations with the default summation
    order, so the tuned sums equal the default ones bit for bit."""
    rng = np.random.default_rng(33)
    x = rng.standard_normal((4096, 64)).astype(np.float32)
    g = A(gpu, x)
    default = {ax: bits(np.asarray(g.sum(axis=ax))) for ax in (0, None)}
    tuned = []
    try:
        for ax in (0, None):
            (r,) = vk.autotune_reduce(gpu, x.shape, ax, reps=2, same_order=True)
            tuned.append(r["shape"])
            cfgs = vk.reduce_configs(r["family"])
            for c, t in zip(cfgs, r["ms"]):
                assert (t > 0) == (c["par"] == cfgs[0]["par"]), (c, t)
            assert cfgs[r["best"]]["par"] == cfgs[0]["par"]
            np.testing.assert_array_equal(bits(np.asarray(g.sum(axis=ax))), default[ax])
    finally:
        for problem in tuned:
            set_cfg(gpu, problem, -1)
