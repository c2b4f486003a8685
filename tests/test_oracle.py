"""Pinning of the oracle (CPU only).

Every oracle function is checked against something other than its own formula:
the reference's known answers (random.py:17-24, test/test_vulkpy.py,
test/test_nn.py), published values (splitmix64), library routines (NumPy
broadcasting, float64 products), brute force on tiny inputs, invariants (GF(2)
linearity, affine invariance) and central finite differences.  Each check is
chosen so that a plausible slip -- a swapped operand, a wrong shift, a dropped
term -- fails it.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

f32 = np.float32


# ---- Xoshiro128++ -------------------------------------------------------------------
def test_reference_docstring_kat():
    # vulkpy/random.py:17-24 (seed 0, 64 states): random(3) then normal(3)
    r = O.Xoshiro128pp(64, 0)
    np.testing.assert_array_equal(r.random(3), np.array([0.42977667, 0.8235899, 0.90622926], dtype=f32))
    np.testing.assert_allclose(r.normal(3), [-2.3403292, 0.7247794, 0.7118352], rtol=1e-6)


def test_splitmix64_published_values():
    # Vigna's splitmix64.c stream from state 0 (x += golden gamma, then mix):
    # the oracle's splitmix64(x) is mix(x + gamma)
    gamma = 0x9E3779B97F4A7C15
    assert O.splitmix64(0) == 0xE220A8397B1DCDAF
    assert O.splitmix64(gamma) == 0x6E789E6AA1B965F4
    assert O.splitmix64((2 * gamma) % 2 ** 64) == 0x06C45D188009454F


def _next_py(s):
    """xoshiro128++ from S. Vigna's published xoshiro128plusplus.c"""
    M = 0xFFFFFFFF
    rotl = lambda x, k: ((x << k) | (x >> (32 - k))) & M
    result = (rotl((s[0] + s[3]) & M, 7) + s[0]) & M
    t = (s[1] << 9) & M
    s[2] ^= s[0]
    s[3] ^= s[1]
    s[1] ^= s[2]
    s[0] ^= s[3]
    s[2] ^= t
    s[3] = rotl(s[3], 11)
    return result


def test_next_matches_published_algorithm():
    rng = np.random.default_rng(0)
    for _ in range(20):
        st = rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
        a, b = st.copy(), [int(v) for v in st]
        for _ in range(50):
            assert O.next_state(a) == _next_py(b)
        assert [int(v) for v in a] == b


def test_seed_state0_is_chained_splitmix():
    # _vkarray.cc:659-663: seed <- splitmix64(seed) four times
    seed = 12345
    words = []
    for _ in range(4):
        seed = O.splitmix64(seed)
        words.append(seed & 0xFFFFFFFF)
    np.testing.assert_array_equal(O.xoshiro_seed(3, 12345)[0], words)


def test_jump_is_gf2_linear_and_not_canonical():
    rng = np.random.default_rng(1)
    for _ in range(5):
        x = rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
        y = rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
        np.testing.assert_array_equal(O.jump_as_written(x ^ y), O.jump_as_written(x) ^ O.jump_as_written(y))
        assert not np.array_equal(O.jump_as_written(x), O.jump_canonical(x))
    # state i = jump(state i-1)
    st = O.xoshiro_seed(4, 99)
    for i in range(1, 4):
        np.testing.assert_array_equal(st[i], O.jump_as_written(st[i - 1]))


def test_canonical_jump_would_change_the_kat():
    # with Vigna's canonical jump the second and third outputs differ from the
    # reference's docstring, which is why the as-written jump is reproduced
    st = O.xoshiro_seed(64, 0)
    s1 = O.jump_canonical(st[0])
    s2 = O.jump_canonical(s1)
    u = [O.next_state(st[0].copy()), O.next_state(s1.copy()), O.next_state(s2.copy())]
    unit = [((v >> 9) | 0x3F800000) for v in u]
    vals = np.array(unit, dtype=np.uint32).view(np.float32) - f32(1)
    np.testing.assert_allclose(vals, [0.42977667, 0.22850382, 0.13257766], rtol=1e-7)


@pytest.mark.parametrize("size,n", [(5, 1), (5, 5), (5, 13), (64, 200), (7, 70)])
def test_chunk_mapping_brute_force(size, n):
    # output k = next() of state k mod size at its (k div size)-th step
    r = O.Xoshiro128pp(size, 3)
    states = r.state.copy()
    got = r.randint(n)
    want = []
    steps = [states[i].copy() for i in range(size)]
    for k in range(n):
        want.append(O.next_state(steps[k % size]))
    np.testing.assert_array_equal(got, want)
    # states >= n (n < size) are not advanced
    for i in range(size):
        np.testing.assert_array_equal(r.state[i], steps[i])


def test_float_conversion_closed_form():
    r = O.Xoshiro128pp(8, 5)
    s = r.state.copy()
    u = r.random(8)
    raw = np.array([O.next_state(s[i]) for i in range(8)], dtype=np.uint64)
    np.testing.assert_array_equal(u, (raw >> 9).astype(np.float64) * 2.0 ** -23)
    assert (u >= 0).all() and (u < 1).all()


def test_box_muller_closed_form_and_invariance():
    rng = np.random.default_rng(2)
    u = rng.random(1001).astype(f32)
    n = 1001
    got = O.box_muller(np.append(u, f32(0.25)), n, 0.0, 1.0)
    j = np.arange(0, n, 2)
    r = np.sqrt(-2 * np.log(1 - u.astype(np.float64)[j]))
    th = 2 * np.pi * np.append(u, f32(0.25)).astype(np.float64)[j + 1]
    np.testing.assert_allclose(got[j], r * np.sin(th), rtol=2e-6, atol=2e-6)
    np.testing.assert_allclose(got[j[:-1] + 1], (r * np.cos(th))[:-1], rtol=2e-6, atol=2e-6)
    # test_random.py:73-89: (normal(mean=5, stddev=3) - 5) / normal() == 3
    for n in (10, 11):
        a1 = O.Xoshiro128pp(64, 0).normal(n)
        a2 = O.Xoshiro128pp(64, 0).normal(n, 5.0, 3.0)
        np.testing.assert_allclose((a2 - 5) / a1, np.full(n, 3.0), rtol=1e-5)


def test_randrange_kat_and_range():
    # test_random.py:99-103
    np.testing.assert_array_equal(O.Xoshiro128pp(64, 4).randrange(5, 3, 4), [3] * 5)
    v = O.Xoshiro128pp(64, 4).randrange(5000, 10, 20)
    assert v.min() >= 10 and v.max() < 20 and len(set(v.tolist())) == 10


# ---- elementwise ------------------------------------------------------------------------
def _shapes(rng):
    nd = int(rng.integers(1, 6))
    shape = [int(rng.integers(1, 6)) for _ in range(nd)]
    sa = tuple(s if rng.random() < 0.6 else 1 for s in shape)[int(rng.integers(0, nd)):] or (1,)
    sb = tuple(s if rng.random() < 0.6 else 1 for s in shape)[int(rng.integers(0, nd)):] or (1,)
    return sa, sb


def test_broadcast_decode_matches_numpy():
    rng = np.random.default_rng(3)
    ops = {"add": np.add, "sub": np.subtract, "mul": np.multiply, "div": np.divide,
           "max": np.maximum, "min": np.minimum}
    for _ in range(200):
        sa, sb = _shapes(rng)
        a = rng.standard_normal(sa).astype(f32)
        b = rng.standard_normal(sb).astype(f32)
        for op, f in ops.items():
            np.testing.assert_array_equal(O.binary_bcast(op, a, b), f(a, b))


def test_same_shape_and_scalar_ops_match_numpy():
    rng = np.random.default_rng(4)
    a = rng.standard_normal(1000).astype(f32)
    b = rng.standard_normal(1000).astype(f32)
    np.testing.assert_array_equal(O.binary("add", a, b), a + b)
    np.testing.assert_array_equal(O.binary("div", a, b), a / b)
    np.testing.assert_array_equal(O.binary("rsub", a, 2.5), f32(2.5) - a)
    np.testing.assert_array_equal(O.binary("rdiv", a, 2.5), f32(2.5) / a)
    p = np.abs(a) + f32(0.1)
    np.testing.assert_array_equal(O.binary("pow", p, b),
                                  (p.astype(np.float64) ** b.astype(np.float64)).astype(f32))
    # NaN propagation of max/min follows numpy
    x = np.array([np.nan, 1, 2], dtype=f32)
    y = np.array([0, np.nan, 3], dtype=f32)
    np.testing.assert_array_equal(O.binary("max", x, y), np.maximum(x, y))


def test_broadcast_to_matches_numpy():
    rng = np.random.default_rng(5)
    for _ in range(50):
        sa, _ = _shapes(rng)
        target = tuple(int(rng.integers(2, 4)) if s == 1 else s for s in sa)
        a = rng.standard_normal(sa).astype(f32)
        np.testing.assert_array_equal(O.broadcast_to(a, (2,) + target), np.broadcast_to(a, (2,) + target))


# ---- reductions ----------------------------------------------------------------------------
def test_reduce_axis_brute_force_and_numpy():
    rng = np.random.default_rng(6)
    x = rng.standard_normal((3, 5, 4)).astype(f32)
    for axis in range(3):
        got = O.reduce_axis("sum", x, axis)
        moved = np.moveaxis(x, axis, 0)
        acc = np.zeros(moved.shape[1:], dtype=f32)
        for k in range(moved.shape[0]):
            acc = (acc + moved[k]).astype(f32)
        np.testing.assert_array_equal(got, acc)
        np.testing.assert_array_equal(O.reduce_axis("maximum", x, axis), x.max(axis=axis))
        np.testing.assert_array_equal(O.reduce_axis("minimum", x, axis), x.min(axis=axis))
        np.testing.assert_allclose(O.reduce_axis("prod", x, axis), x.astype(np.float64).prod(axis=axis), rtol=1e-6)
        rb = O.reduce_axis("maximum", x, axis, rebroadcast=True)
        np.testing.assert_array_equal(rb, np.broadcast_to(x.max(axis=axis, keepdims=True), x.shape))
    np.testing.assert_allclose(O.reduce_axis("sum", x, (0, 2)), x.sum(axis=(0, 2)), rtol=1e-6)
    # reference KATs test_vulkpy.py:853-869, 1444-1472
    ones = np.ones((2, 3, 4, 2, 2, 4, 3), dtype=f32)
    np.testing.assert_array_equal(O.reduce_axis("sum", ones, (1, 2, 5)), ones.sum(axis=(1, 2, 5)))
    m = np.array([[1, 2], [3, 4]], dtype=f32)
    np.testing.assert_array_equal(O.reduce_axis("prod", m, 0, rebroadcast=True), [[3, 8], [3, 8]])
    np.testing.assert_array_equal(O.mean_axis(m, 0), [2, 3])


def test_sum_f64():
    x = np.random.default_rng(7).standard_normal(100001).astype(f32)
    assert abs(O.sum_f64(x) - math.fsum(x.astype(np.float64).tolist())) < 1e-9


# ---- matmul / affine / gather -------------------------------------------------------------------
def test_matmul_kat_and_float64():
    np.testing.assert_array_equal(O.matmul([[1, 2], [3, 4]], [[1, 2], [3, 4]]), [[7, 10], [15, 22]])
    rng = np.random.default_rng(8)
    a = rng.integers(-8, 9, (17, 23)).astype(f32)
    b = rng.integers(-8, 9, (23, 9)).astype(f32)
    np.testing.assert_array_equal(O.matmul(a, b), a @ b)  # exact in FP32
    a = rng.standard_normal((40, 300)).astype(f32)
    b = rng.standard_normal((300, 30)).astype(f32)
    ref = a.astype(np.float64) @ b.astype(np.float64)
    assert np.abs(O.matmul(a, b) - ref).max() < 300 * 2 ** -23 * np.abs(a).max() * np.abs(b).max()


def test_batch_affine():
    rng = np.random.default_rng(9)
    W = rng.integers(-4, 5, (6, 5)).astype(f32)
    b = rng.integers(-4, 5, 6).astype(f32)
    x = rng.integers(-4, 5, (3, 5)).astype(f32)
    np.testing.assert_array_equal(O.batch_affine(W, b, x), x @ W.T + b)


def test_gather_kats():
    a = np.arange(8).reshape(2, 2, 2)
    idx = np.array([[0, 1], [6, 7]])
    np.testing.assert_array_equal(O.gather(a, idx), np.ravel(a)[idx])
    a = np.arange(30).reshape(3, 2, 5)
    idx = np.array([[0, 1], [1, 1]])
    got = O.gather_axis(a, idx, 1)
    np.testing.assert_array_equal(got, np.moveaxis(np.take(a, idx, axis=1), (1, 2), (0, 1)))
    np.testing.assert_array_equal(O.gather_axis(np.identity(3), [0, 1, 2, 1, 0], 0),
                                  np.identity(3)[[0, 1, 2, 1, 0]])


# ---- nn restatements --------------------------------------------------------------------------
def test_activation_kats():
    # test_nn.py:70-153
    np.testing.assert_array_equal(O.relu([[-0.2, 0.0, 0.2]]), np.array([[0, 0, 0.2]], dtype=f32))
    np.testing.assert_array_equal(O.relu_backward(O.relu([[-0.2, 0.0, 0.2]]), [[0.7, 0.8, 0.9]]),
                                  np.array([[0, 0, 0.9]], dtype=f32))
    d = np.array([[-100, -0.1, 0, 10, 100]])
    np.testing.assert_allclose(O.sigmoid(d), 1 / (1 + np.exp(-d)), rtol=1e-7, atol=1e-7)
    np.testing.assert_array_equal(O.softmax([[1.0, 1.0]]), [[0.5, 0.5]])
    np.testing.assert_array_equal(O.softmax([[100.0, 0.0]]), [[1.0, 0.0]])
    e = np.exp(d - d.max(axis=1))
    np.testing.assert_allclose(O.softmax(d), e / e.sum(axis=1, keepdims=True), rtol=1e-7, atol=1e-7)
    y = O.sigmoid(d)
    dy = np.array([[0.1, 0.2, 0.3, 0.5, 0.7]])
    np.testing.assert_allclose(O.sigmoid_backward(y, dy), dy * y * (1 - y), rtol=1e-6)


def test_dense_backward_kat_and_finite_differences():
    # test_nn.py:181-200
    x = np.array([[1, 2], [3, 4]], dtype=f32)
    dy = np.array([[4, 2], [1, 3]], dtype=f32)
    W = np.array([[0.5, -1.0], [2.0, 0.25]], dtype=f32)
    dW, db, dx = O.dense_backward_reference(x, dy, W)
    np.testing.assert_array_equal(dW, [[7, 12], [11, 16]])
    np.testing.assert_array_equal(db, [5, 5])
    np.testing.assert_allclose(dx, dy @ W)
    # FP64 central differences of L = sum(dy * (x W^T + b))
    rng = np.random.default_rng(10)
    x = rng.standard_normal((8, 5))
    W = rng.standard_normal((4, 5))
    b = rng.standard_normal(4)
    dy = rng.standard_normal((8, 4))
    L = lambda W_, b_, x_: float(np.sum(dy * (x_ @ W_.T + b_)))
    h = 1e-6
    gW = np.zeros_like(W)
    for i in range(4):
        for j in range(5):
            E = np.zeros_like(W); E[i, j] = h
            gW[i, j] = (L(W + E, b, x) - L(W - E, b, x)) / (2 * h)
    gb = np.array([(L(W, b + h * np.eye(4)[i], x) - L(W, b - h * np.eye(4)[i], x)) / (2 * h) for i in range(4)])
    gx = np.zeros_like(x)
    for i in range(8):
        for j in range(5):
            E = np.zeros_like(x); E[i, j] = h
            gx[i, j] = (L(W, b, x + E) - L(W, b, x - E)) / (2 * h)
    dW, db, dx = O.dense_backward_reference(x, dy, W)
    np.testing.assert_allclose(dW, gW, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(db, gb, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(dx, gx, rtol=1e-5, atol=1e-5)


def test_loss_kats():
    # test_nn.py:207-223
    assert O.batch_mean(O.cross_entropy_rows([[1.0, 0.0]], [[1.0, 0.0]])) == 0.0
    np.testing.assert_allclose(O.batch_mean(O.cross_entropy_rows([[0.5, 0.5]], [[0.5, 0.5]])), 0.6931472, rtol=1e-7)
    _x = np.asarray([[0.7, 0.3], [0.2, 0.8], [1.0, 0.0]])
    _y = np.asarray([[1.0, 0.0], [0.0, 1.0], [1.0, 0.0]])
    _L = np.sum(-_y * np.log(_x + 1e-8), axis=1)
    np.testing.assert_allclose(O.batch_mean(O.cross_entropy_rows(_x, _y)), _L.mean(), rtol=1e-7, atol=1e-7)
    np.testing.assert_allclose(O.cross_entropy_backward(_x, _y, 1 / 3), -_y / (_x + 1e-8) / 3, rtol=1e-6)
    # softmax cross entropy, test_nn.py:279-363
    _x = np.asarray([[-1, 0], [10, 15]], dtype=f32)
    _y = np.asarray([[1, 0], [0, 1]], dtype=f32)
    e = np.exp(_x - _x.max(axis=1, keepdims=True))
    sm = e / e.sum(axis=1, keepdims=True)
    np.testing.assert_allclose(O.batch_mean(O.cross_entropy_rows(O.softmax(_x), _y)),
                               (-_y * np.log(sm)).sum(axis=1).mean(), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(O.loss_grad("sce", O.softmax(_x), _y, 0.5), (sm - _y) / 2, atol=1e-7)
    # MSE, test_nn.py:365-414
    _x = np.asarray([[4, 2], [1, 1.5]])
    _y = np.asarray([[3, 2.2], [0.7, 1.5]])
    np.testing.assert_allclose(O.batch_mean(O.mse_rows(_x, _y)), np.square(_y - _x).sum(axis=1).mean(), rtol=1e-6)
    np.testing.assert_allclose(O.loss_grad("mse", _x, _y, 0.5), _x - _y, rtol=1e-6)
    np.testing.assert_allclose(O.loss_grad("mse", _x, _y), 2 * (_x - _y), rtol=1e-6)
    # Huber, test_nn.py:416-478
    _x = np.asarray([[1.0, 2.2], [-3.0, 0.7]])
    _y = np.asarray([[10, 3.0], [-5, 0.5]])
    np.testing.assert_allclose(O.batch_mean(O.huber_rows(_x, _y)), 2.92, rtol=1e-6)
    np.testing.assert_allclose(O.huber_rows(_x, _y).sum(), 5.84, rtol=1e-6)
    np.testing.assert_allclose(O.loss_grad("huber", _x, _y, 0.5), [[-0.5, -0.4], [0.5, 0.1]], rtol=1e-6)
    np.testing.assert_allclose(O.loss_grad("huber", _x, _y), [[-1.0, -0.8], [1.0, 0.2]], rtol=1e-6)


def test_regularizer_grad_kats_and_closed_form():
    # test/test_nn.py:487-549: W = 0, 3.5, -3.5 with coefficients 1
    for v, ridge, lasso, elastic in ((0.0, 0.0, 0.0, 0.0), (3.5, 7.0, 1.0, 8.0), (-3.5, -7.0, -1.0, -8.0)):
        w = np.float32([v])
        assert O.regularizer_grad("ridge", w, l2=1.0)[0] == ridge
        assert O.regularizer_grad("lasso", w, l1=1.0)[0] == lasso
        assert O.regularizer_grad("elastic", w, l1=1.0, l2=1.0)[0] == elastic
    # closed form in FP64 for other coefficients: each term is one rounding,
    # so a dropped factor 2 or a swapped coefficient shows at once
    rng = np.random.default_rng(5)
    w = rng.standard_normal(10000).astype(np.float32)
    w[:3] = [0.0, -0.0, 1e-42]
    c1, c2 = 3e-4, 7e-3
    r = O.regularizer_grad("ridge", w, l2=c2).astype(np.float64)
    exact = 2 * c2 * w.astype(np.float64)
    assert np.all(np.abs(r - exact) <= 2 * np.spacing(np.abs(exact).astype(np.float32)) + 1e-45)
    assert np.array_equal(O.regularizer_grad("lasso", w, l1=c1), np.float32(c1) * np.sign(w))
    e = O.regularizer_grad("elastic", w, l1=c1, l2=c2).astype(np.float64)
    ex = c1 * np.sign(w) + 2 * c2 * w.astype(np.float64)
    assert np.all(np.abs(e - ex) <= 4 * np.spacing(np.abs(ex).astype(np.float32)) + 1e-45)


def test_optimizers_against_textbook_formulas():
    rng = np.random.default_rng(11)
    w = rng.standard_normal(100).astype(f32)
    m = np.zeros(100, f32)
    v = np.zeros(100, f32)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    w64, m64, v64 = w.astype(np.float64), np.zeros(100), np.zeros(100)
    b1t = b2t = 1.0
    for t in range(1, 6):
        g = rng.standard_normal(100).astype(f32)
        b1t *= b1
        b2t *= b2
        w, m, v = O.adam(w, g, m, v, lr, b1, b2, eps, b1t, b2t)
        g64 = g.astype(np.float64)
        m64 = b1 * m64 + (1 - b1) * g64
        v64 = b2 * v64 + (1 - b2) * g64 ** 2
        w64 = w64 - lr * (m64 / (1 - b1 ** t)) / (np.sqrt(v64 / (1 - b2 ** t)) + eps)
    np.testing.assert_allclose(w, w64, rtol=1e-6, atol=1e-7)
    g = rng.standard_normal(100).astype(f32)
    np.testing.assert_allclose(O.sgd(w, g, 0.01), w - 0.01 * g, rtol=1e-6)
    h = np.full(100, 0.5, f32)
    w2, h2 = O.adagrad(w, g, h, 0.01, 1e-8)
    np.testing.assert_allclose(h2, 0.5 + g.astype(np.float64) ** 2, rtol=1e-6)
    np.testing.assert_allclose(w2, w - 0.01 * g / (np.sqrt(h2) + 1e-8), rtol=1e-6, atol=1e-7)


def test_fastdiv_header_on_host(tmp_path):
    # vulkpy_amd/csrc/fastdiv.h computes the row/column of every vectorised
    # elementwise unit; a wrong quotient would index out of bounds
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "fastdiv_check"
    subprocess.run(["hipcc", "-O2", "-std=c++17", os.path.join(root, "tests", "fastdiv_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert out[0] == "0" and int(out[1]) > 10 ** 8


def test_unary_header_on_host(tmp_path):
    # tanh / log / log2 / pow / sinh / cosh / asinh / acosh / atanh / sin /
    # cos / tan / asin / acos / atan / exp / exp2 / invsqrt of
    # vulkpy_amd/csrc/bmath.h (the elementwise kernels' FP64 evaluations)
    # against the FP64 library functions: every 997th FP32 input of each
    # computing range (pow: every 63811th positive base against 31 exponents)
    # within 0.5 ulp + 2^-13 of the exact value, and the special values
    # (signed zeros, negatives, infinities, NaN, 1, powers of two, domain
    # edges) as the FP64 functions rounded
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "unary_check"
    subprocess.run(["hipcc", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(root, "tests", "unary_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe), "997"], check=True, capture_output=True, text=True).stdout.split()
    assert len(out) == 20, out
    assert all(float(e) <= 0.5 + 2 ** -13 for e in out[:18]), out
    assert int(out[18]) > 3 * 10 ** 7 and out[19] == "0", out


def test_sincos_header_on_host(tmp_path):
    # vulkpy_amd/csrc/bmath.h gives the fused normal kernel its sin / cos; it
    # is checked at every angle the generator can produce (2^23 of them)
    # against FP64 sin / cos: within 2 ulp of the correctly rounded value
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "sincos_check"
    subprocess.run(["hipcc", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(root, "tests", "sincos_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert float(out[0]) <= 2.0 and float(out[1]) <= 2.0 and int(out[2]) == 1 << 23, out
