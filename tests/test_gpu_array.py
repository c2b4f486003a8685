"""GPU parity of the array path (F1-F17) against the oracle and the reference's
own known answers (test/test_vulkpy.py).  Integer-exact ops (+ - * /, max,
min, clamp, copies, gather) are compared bit for bit; transcendentals against
FP64 with ulp bounds; reductions against the oracle's sequential order or
FP64 with stated tolerances."""
import numpy as np
import pytest

import vulkpy_amd as vk
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ULP32 = np.float32(2.0 ** -23)


def A(gpu, x):
    return vk.Array(gpu, data=np.asarray(x, dtype=np.float32))


def assert_bits(got, want):
    got = np.asarray(got, dtype=np.float32)
    want = np.asarray(want, dtype=np.float32)
    assert got.shape == want.shape, (got.shape, want.shape)
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert same.all(), f"{(~same).sum()} mismatches, first at {np.argwhere(~same)[:3].tolist()}"


def ulp_err(got, exact64):
    got = np.asarray(got, dtype=np.float64)
    exact64 = np.asarray(exact64, dtype=np.float64)
    spacing = np.spacing(np.abs(exact64).astype(np.float32)).astype(np.float64)
    return np.abs(got - exact64) / spacing


# ---- reference known answers (test/test_vulkpy.py) -------------------------------
def test_reference_arithmetic_kats(gpu):
    a, b = A(gpu, [5, 5, 5]), A(gpu, [1, 1, 1])
    np.testing.assert_allclose(a + b, [6, 6, 6])
    c = A(gpu, [4, 4, 4])
    np.testing.assert_allclose(a + b + c, [10, 10, 10])
    np.testing.assert_allclose(A(gpu, [4, 4, 4]) - A(gpu, [2, 2, 2]), [2, 2, 2])
    np.testing.assert_allclose(A(gpu, [2, 2, 2]) * A(gpu, [3, 3, 3]), [6, 6, 6])
    np.testing.assert_allclose(A(gpu, [8, 8, 8]) / A(gpu, [2, 2, 2]), [4, 4, 4])
    x = A(gpu, [1, 2, 3])
    np.testing.assert_allclose(x + 1, [2, 3, 4])
    np.testing.assert_allclose(1 + x, [2, 3, 4])
    np.testing.assert_allclose(5 - x, [4, 3, 2])
    np.testing.assert_allclose(6 / x, [6, 3, 2])
    y = A(gpu, [1, 2, 3])
    y += 1
    y *= 2
    y -= 1
    y /= 2
    np.testing.assert_allclose(y, [1.5, 2.5, 3.5])
    with pytest.raises(ValueError):
        A(gpu, [1, 1]) + A(gpu, [1, 1, 1])


def test_reference_matmul_kats(gpu):
    a = A(gpu, [[1, 2], [3, 4]])
    np.testing.assert_allclose(a @ a, [[7, 10], [15, 22]])
    np.testing.assert_allclose(a @ A(gpu, [1, 3]), [7, 15])
    np.testing.assert_allclose(A(gpu, [1, 2]) @ a, [7, 10])
    with pytest.raises(ValueError):
        A(gpu, [[1, 2, 3], [4, 5, 6]]) @ A(gpu, [[1, 2, 3], [4, 5, 6]])
    v = A(gpu, [1, 2, 3]) @ A(gpu, [4, 5, 6])
    assert v.shape == (1,)
    np.testing.assert_allclose(v, [32])


def test_reference_reshape_and_minmax(gpu):
    a = A(gpu, [1, 2, 3, 4])
    a.reshape((2, 2))
    np.testing.assert_allclose(a, [[1, 2], [3, 4]])
    with pytest.raises(ValueError):
        A(gpu, [1, 2, 3]).reshape((2, 2))
    a, b = A(gpu, [1, 2, 3]), A(gpu, [3, 2, 1])
    np.testing.assert_allclose(a.max(b), [3, 2, 3])
    np.testing.assert_allclose(a.min(b), [1, 2, 1])
    np.testing.assert_allclose(a.max(1.5), [1.5, 2, 3])
    np.testing.assert_allclose(a.min(1.5), [1, 1.5, 1.5])
    a.max(b, inplace=True)
    np.testing.assert_allclose(a, [3, 2, 3])
    c = A(gpu, [1, 2, 3])
    c.min(1.5, inplace=True)
    np.testing.assert_allclose(c, [1, 1.5, 1.5])


def test_reference_broadcast_kats(gpu):
    one21, one22 = A(gpu, np.ones((2, 1))), A(gpu, np.ones((2, 2)))
    np.testing.assert_allclose(one21 + one22, [[2, 2], [2, 2]])
    np.testing.assert_allclose(A(gpu, np.ones(2)) + one22, [[2, 2], [2, 2]])
    _a = np.arange(4).reshape((1, 2, 2))
    _b = np.arange(4).reshape((2, 2, 1))
    np.testing.assert_allclose(A(gpu, _a) + A(gpu, _b), _a + _b)
    m = A(gpu, [[1, 2], [3, 4]])
    np.testing.assert_allclose(m.max(A(gpu, [[2, 3]])), [[2, 3], [3, 4]])
    np.testing.assert_allclose(m.min(A(gpu, [2, 3])), [[1, 2], [2, 3]])
    np.testing.assert_allclose(m ** A(gpu, [[2, 3]]), [[1, 8], [9, 64]])
    i = A(gpu, np.ones((2, 2)))
    i += A(gpu, np.ones((2, 1)))
    np.testing.assert_allclose(i, 2 * np.ones((2, 2)))
    i = A(gpu, 2 * np.ones((2, 2)))
    i /= A(gpu, np.ones(2))
    np.testing.assert_allclose(i, 2 * np.ones((2, 2)))
    b = A(gpu, np.ones((1, 2, 2))).broadcast_to((3, 2, 2))
    np.testing.assert_allclose(b, np.ones((3, 2, 2)))
    np.testing.assert_allclose(A(gpu, [1, 2, 3]).broadcast_to((2, 3)), [[1, 2, 3], [1, 2, 3]])
    with pytest.raises(ValueError):
        A(gpu, np.ones(3)).broadcast_to((4,))


def test_reference_clamp_kats(gpu):
    x = np.asarray([1.2, 3.5, 10])
    lo, hi = np.asarray([3, 3, 4]), np.asarray([6, 6, 7])
    np.testing.assert_allclose(A(gpu, x).clamp(A(gpu, lo), A(gpu, hi)), np.clip(x, lo, hi))
    np.testing.assert_allclose(A(gpu, x).clamp(3, 7), np.clip(x, 3, 7))
    a = A(gpu, x)
    a.clamp(A(gpu, lo), 7, inplace=True)
    np.testing.assert_allclose(a, np.clip(x, lo, 7))
    m = A(gpu, [[1, 2], [3, 4]])
    b, c = A(gpu, [[2, 3]]), A(gpu, [2, 3])
    np.testing.assert_allclose(m.clamp(b, c), [[2, 3], [2, 3]])
    np.testing.assert_allclose(m.clamp(b, 5), [[2, 3], [3, 4]])
    np.testing.assert_allclose(m.clamp(2, c), [[2, 2], [2, 3]])
    m.clamp(2, c, inplace=True)
    np.testing.assert_allclose(m, [[2, 2], [2, 3]])


def test_reference_reduction_kats(gpu):
    a = A(gpu, [1, 2, 3])
    np.testing.assert_allclose(a.sum(), (6,))
    np.testing.assert_allclose(a.prod(), (6,))
    np.testing.assert_allclose(a.maximum(), (3,))
    np.testing.assert_allclose(a.minimum(), (1,))
    np.testing.assert_allclose(a.mean(), (2,))
    np.testing.assert_allclose(a.sum(axis=0), (6,))
    ones = A(gpu, np.ones((2, 2)))
    np.testing.assert_allclose(ones.sum(keepdims=True), [[4]])
    np.testing.assert_allclose(ones.sum(axis=0, keepdims=True), [[2, 2]])
    big = A(gpu, np.ones(65))
    for f, v in (("sum", 65), ("prod", 1), ("maximum", 1), ("minimum", 1), ("mean", 1)):
        np.testing.assert_allclose(getattr(big, f)(), (v,))
    x = np.ones((2, 3, 4, 2, 2, 4, 3))
    g = A(gpu, x)
    np.testing.assert_allclose(g.sum(axis=(1, 2, 5)), x.sum(axis=(1, 2, 5)))
    np.testing.assert_allclose(g.mean(axis=(1, 2, 5)), x.mean(axis=(1, 2, 5)))
    m = A(gpu, [[1, 2], [3, 4]])
    np.testing.assert_allclose(m.sum(axis=0, rebroadcast=True), [[4, 6], [4, 6]])
    np.testing.assert_allclose(m.prod(axis=0, rebroadcast=True), [[3, 8], [3, 8]])
    np.testing.assert_allclose(m.maximum(axis=0, rebroadcast=True), [[3, 4], [3, 4]])
    np.testing.assert_allclose(m.minimum(axis=0, rebroadcast=True), [[1, 2], [1, 2]])
    np.testing.assert_allclose(m.mean(axis=0, rebroadcast=True), [[2, 3], [2, 3]])
    with pytest.raises(ValueError):
        m.sum(axis=(0, 1), rebroadcast=True)


def test_reference_gather_kats(gpu):
    _a = np.arange(8).reshape((2, 2, 2))
    for _idx in (np.asarray([0, 1, 6, 7]), np.asarray([[0, 1], [6, 7]])):
        b = A(gpu, _a).gather(vk.U32Array(gpu, data=_idx))
        assert tuple(b.shape) == _idx.shape
        np.testing.assert_allclose(b, np.ravel(_a)[_idx])
    _a = np.arange(30).reshape((3, 2, 5))
    _idx = np.asarray([[0, 1], [1, 1]])
    b = A(gpu, _a).gather(vk.U32Array(gpu, data=_idx), axis=1)
    assert tuple(b.shape) == (2, 2, 3, 5)
    np.testing.assert_allclose(b, np.moveaxis(np.take(_a, _idx, axis=1), (1, 2), (0, 1)))
    oh = vk.U32Array(gpu, data=[0, 1, 2, 1, 0]).to_onehot(3)
    np.testing.assert_allclose(oh, np.identity(3)[[0, 1, 2, 1, 0]])
    oh = vk.U32Array(gpu, data=[[0, 1], [2, 1]]).to_onehot(3)
    np.testing.assert_allclose(oh, np.identity(3)[[[0, 1], [2, 1]]])
    # labels past the last class take the last class, as the reference's
    # gather of identity rows clamps them; large label sets bit for bit
    lab = np.asarray([0, 5, 2, 3, 1], dtype=np.uint32)
    np.testing.assert_array_equal(vk.U32Array(gpu, data=lab).to_onehot(3),
                                  np.identity(3, dtype=np.float32)[np.minimum(lab, 2)])
    lab = np.random.default_rng(4).integers(0, 10, 4096).astype(np.uint32)
    got = vk.U32Array(gpu, data=lab).to_onehot(10)
    assert got.shape == (4096, 10)
    np.testing.assert_array_equal(got, np.identity(10, dtype=np.float32)[lab])


def test_reference_math_kats(gpu):
    x = np.asarray([1, 2, 3])
    for f, ref, tol in (("exp", np.exp, 1e-7), ("log", np.log, 1e-7), ("exp2", np.exp2, 1e-7),
                        ("log2", np.log2, 1e-7), ("sqrt", np.sqrt, 1e-7),
                        ("invsqrt", lambda v: np.sqrt(1 / v), 1e-7)):
        np.testing.assert_allclose(getattr(A(gpu, x), f)(), ref(x), rtol=tol)
        a = A(gpu, x)
        getattr(a, f)(inplace=True)
        np.testing.assert_allclose(a, ref(x), rtol=tol)
    t = np.asarray([-0.9, 0.1, 0.5])
    trig = {"sin": np.sin, "cos": np.cos, "tan": np.tan, "atan": np.arctan, "sinh": np.sinh,
            "cosh": np.cosh, "tanh": np.tanh, "asinh": np.arcsinh, "atanh": np.arctanh}
    for f, npf in trig.items():
        np.testing.assert_allclose(getattr(A(gpu, t), f)(), npf(t), rtol=1e-5)
    for f, npf in (("asin", np.arcsin), ("acos", np.arccos)):
        np.testing.assert_allclose(getattr(A(gpu, t), f)(), npf(t), rtol=1e-3)
    np.testing.assert_allclose(A(gpu, [1.5, 2.5]).acosh(), np.arccosh([1.5, 2.5]), rtol=1e-5)
    np.testing.assert_allclose(A(gpu, [-1, 1, 3]).abs(), [1, 1, 3])
    np.testing.assert_allclose(A(gpu, [-1, 0, 3]).sign(), [-1, 0, 1])
    y = np.asarray([1.1, 2.2, 1.4])
    np.testing.assert_allclose(A(gpu, x) ** A(gpu, y), x ** y, rtol=1e-7)
    np.testing.assert_allclose(A(gpu, x) ** 2.7, x ** 2.7, rtol=1e-6)
    np.testing.assert_allclose(1.3 ** A(gpu, y), 1.3 ** y, rtol=1e-6)


def test_reference_inplace_kats(gpu):
    """The in-place and ordering tests of test/test_vulkpy.py: every unary
    function in place (inputs and tolerances of test_i<fn>), **= with an
    array, a scalar and broadcast operands, in-place broadcast max / min /
    arithmetic, clamp in place with every bound kind, and the cascade /
    chain / unrelated-buffer / write-then-op / 2-D cases."""
    cases = {
        "abs": ([-1, 1, 3], lambda v: np.abs(v), 1e-7), "sign": ([-1, 1, 3], np.sign, 1e-7),
        "sin": ([1, 3, 5], np.sin, 1e-5), "cos": ([1, 3, 5], np.cos, 1e-5),
        "tan": ([1, 3, 5], np.tan, 1e-5), "sinh": ([1, 3, 5], np.sinh, 1e-5),
        "cosh": ([1, 3, 5], np.cosh, 1e-5), "tanh": ([1, 3, 5], np.tanh, 1e-5),
        "asin": ([0.5, 0.3, -0.2], np.arcsin, 1e-3), "acos": ([0.5, 0.3, -0.2], np.arccos, 1e-3),
        "atan": ([0.5, 0.3, -0.2], np.arctan, 1e-5), "asinh": ([0.5, 0.3, -0.2], np.arcsinh, 1e-5),
        "acosh": ([1.5, 1.3, 1.2], np.arccosh, 1e-5), "atanh": ([0.5, 0.3, -0.2], np.arctanh, 1e-5),
        "exp": ([1, 2, 3], np.exp, 1e-7), "log": ([1, 2, 3], np.log, 1e-7),
        "exp2": ([1, 2, 3], np.exp2, 1e-7), "log2": ([1, 2, 3], np.log2, 1e-7),
        "sqrt": ([1, 2, 3], np.sqrt, 1e-7), "invsqrt": ([1, 2, 3], lambda v: np.sqrt(1 / v), 1e-7),
    }
    for f, (x, ref, rtol) in cases.items():
        a = A(gpu, x)
        out = getattr(a, f)(inplace=True)
        a.wait()
        np.testing.assert_allclose(a, ref(np.asarray(x, dtype=np.float64)), rtol=rtol, err_msg=f)
        assert out is a or out is None, f
    x, y = np.asarray([1, 2, 3]), np.asarray([1.1, 2.2, 1.4])
    a = A(gpu, x)
    a **= A(gpu, y)
    np.testing.assert_allclose(a, x ** y, rtol=1e-7)
    a = A(gpu, x)
    a **= 2.7
    np.testing.assert_allclose(a, x ** 2.7, rtol=1e-6)
    for b in ([[2, 3]], [2, 3]):                    # test_ipow_broadcast(_newaxis)
        a = A(gpu, [[1, 2], [3, 4]])
        a **= A(gpu, b)
        np.testing.assert_allclose(a, [[1, 8], [9, 64]])
    for b in ([[2, 3]], [2, 3]):                    # test_imax/imin_broadcast(_newaxis)
        a = A(gpu, [[1, 2], [3, 4]])
        a.max(A(gpu, b), inplace=True)
        np.testing.assert_allclose(a, [[2, 3], [3, 4]])
        a = A(gpu, [[1, 2], [3, 4]])
        a.min(A(gpu, b), inplace=True)
        np.testing.assert_allclose(a, [[1, 2], [2, 3]])
    ones = np.ones((2, 2))
    for op, b, want in (("__isub__", np.ones(2), np.zeros((2, 2))), ("__imul__", np.ones((2, 1)), ones),
                        ("__iadd__", np.ones(2), 2 * ones), ("__itruediv__", np.ones((2, 1)), ones)):
        a = A(gpu, ones)
        a = getattr(a, op)(A(gpu, b))
        np.testing.assert_allclose(a, want, err_msg=op)
    x, lo, hi = np.asarray([1.2, 3.5, 10]), np.asarray([3, 3, 4]), np.asarray([6, 6, 7])
    for lo_, hi_ in ((lo, hi), (3, hi), (lo, 7), (3, 7)):   # test_iclamp, _sv, _vs, _ss
        a = A(gpu, x)
        a.clamp(A(gpu, lo_) if isinstance(lo_, np.ndarray) else lo_,
                A(gpu, hi_) if isinstance(hi_, np.ndarray) else hi_, inplace=True)
        np.testing.assert_allclose(a, np.clip(x, lo_, hi_))
    m = A(gpu, [[1, 2], [3, 4]])
    m.clamp(A(gpu, [[2, 3]]), A(gpu, [2, 3]), inplace=True)  # test_iclamp_broadcast
    np.testing.assert_allclose(m, [[2, 3], [2, 3]])
    # ordering: cascade, unrelated buffers, chain, write then op, 2-D
    a, b, c = A(gpu, [5, 5, 5]), A(gpu, [1, 1, 1]), A(gpu, [4, 4, 4])
    d = a + b + c
    d.wait()
    np.testing.assert_allclose(d, [10, 10, 10])
    e, f = A(gpu, [5, 5, 5]) + A(gpu, [1, 1, 1]), A(gpu, [2, 2, 2]) + A(gpu, [4, 4, 4])
    np.testing.assert_allclose(e, f)
    a, b = A(gpu, [1, 1, 1]), A(gpu, [2, 2, 2])
    c = a + b
    d = a + c
    d.wait()
    np.testing.assert_allclose(d, [4, 4, 4])
    a, b = vk.Array(gpu, shape=(3,)), vk.Array(gpu, shape=(3,))
    a[:] = 5
    b[:] = 7
    np.testing.assert_allclose(a + b, [12, 12, 12])
    np.testing.assert_allclose(A(gpu, [[1, 1], [1, 1]]) + A(gpu, [[2, 2], [2, 2]]), [[3, 3], [3, 3]])
    np.testing.assert_allclose(A(gpu, np.ones((2, 2))).broadcast_to((3, 2, 2)), np.ones((3, 2, 2)))


# ---- oracle parity on the C1 workload ---------------------------------------------
@pytest.fixture(scope="module")
def c1():
    rng = np.random.default_rng(0)
    A_ = rng.uniform(-1, 1, (1024, 1024)).astype(np.float32)
    B_ = rng.uniform(-1, 1, (1024, 1024)).astype(np.float32)
    return A_, B_


def test_c1_exact_ops(gpu, c1):
    a, b = c1
    ga, gb = A(gpu, a), A(gpu, b)
    assert_bits(ga + gb, a + b)
    assert_bits(ga * gb, a * b)
    assert_bits(ga / (gb.abs() + 1), a / (np.abs(b) + np.float32(1)))
    assert_bits(ga.clamp(-0.5, 0.5), np.clip(a, np.float32(-0.5), np.float32(0.5)))
    assert_bits(ga - gb, a - b)
    assert_bits(ga.max(gb), np.maximum(a, b))
    assert_bits(ga.min(gb), np.minimum(a, b))
    assert_bits(ga.sqrt(), np.sqrt(a))  # NaN for negative inputs on both sides


def test_c1_transcendentals(gpu, c1):
    a, b = c1
    ga, gb = A(gpu, a), A(gpu, b)
    a64 = a.astype(np.float64)
    assert ulp_err(ga.exp(), np.exp(a64)).max() <= 0.51
    assert ulp_err(ga.tanh(), np.tanh(a64)).max() <= 0.51
    p = (np.abs(a64) + np.float64(np.float32(0.1))) ** b.astype(np.float64)
    got = (ga.abs() + 0.1) ** gb
    base = (np.abs(a) + np.float32(0.1)).astype(np.float64)
    assert ulp_err(got, base ** b.astype(np.float64)).max() <= 0.51
    assert np.allclose(np.asarray(got), p, rtol=1e-6)


def _pow_inputs():
    rng = np.random.default_rng(41)
    bits = rng.integers(0, 2 ** 32, 1 << 18, dtype=np.uint64).astype(np.uint32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 2.0, 0.5, 1e-45, -1e-45,
                        1.17549435e-38, 3.4028235e38, -3.4028235e38, 1.5, -2.5], dtype=np.float32)
    return np.concatenate([special, bits.view(np.float32)])


@pytest.mark.parametrize("s", [1.0, 2.0, -1.0, 0.0, -0.0, 0.5])
def test_scalar_pow_fast_paths(gpu, s):
    """x ** s with a scalar s in {1, 2, -1, 0, 1/2} is one correctly rounded
    operation: bit for bit the FP64 pow rounded once (double rounding is
    innocuous for these), over every kind of FP32 input (zeros, infinities,
    NaN, denormals, negatives); in place too.  The general kernel is within
    its 0.51-ulp bar of the same values."""
    x = _pow_inputs()
    with np.errstate(all="ignore"):
        exact = x.astype(np.float64) ** np.float64(s)
    if s == 0.5:
        # NumPy takes sqrt for ** 0.5; C pow(-0, 0.5) = +0 and pow(-inf, 0.5) = +inf
        exact[x == 0] = 0.0
        exact[x == -np.inf] = np.inf
    want = exact.astype(np.float32)
    gx = A(gpu, x)
    assert_bits(gx ** s, want)
    y = gx * 1.0
    y **= s
    assert_bits(y, want)
    general = np.asarray(gx ** A(gpu, np.full(x.shape, s, dtype=np.float32)))
    fin = np.isfinite(want) & (want != 0)
    assert ulp_err(general[fin], exact[fin]).max() <= 0.51
    assert_bits(general[~fin], want[~fin])


@pytest.mark.parametrize("s", [1.5, -2.0])
def test_scalar_pow_general_equals_array_exponent(gpu, s):
    x = _pow_inputs()
    gx = A(gpu, x)
    assert_bits(gx ** s, gx ** A(gpu, np.full(x.shape, s, dtype=np.float32)))


# every unary function: the domain each computes on, the input's map from the
# C1 operand A ~ U[-1, 1), and the FP64 reference
UNARY = {
    "abs": (lambda a: a * 8, np.abs), "sign": (lambda a: a, np.sign),
    "sin": (lambda a: a * 8, np.sin), "cos": (lambda a: a * 8, np.cos),
    "tan": (lambda a: a * 8, np.tan), "asin": (lambda a: a, np.arcsin),
    "acos": (lambda a: a, np.arccos), "atan": (lambda a: a * 8, np.arctan),
    "sinh": (lambda a: a * 8, np.sinh), "cosh": (lambda a: a * 8, np.cosh),
    "tanh": (lambda a: a * 8, np.tanh), "asinh": (lambda a: a * 8, np.arcsinh),
    "acosh": (lambda a: 1 + np.abs(a) * 8, np.arccosh), "atanh": (lambda a: a, np.arctanh),
    "exp": (lambda a: a * 8, np.exp), "log": (lambda a: np.abs(a) * 8, np.log),
    "exp2": (lambda a: a * 8, np.exp2), "log2": (lambda a: np.abs(a) * 8, np.log2),
    "sqrt": (lambda a: np.abs(a) * 8, np.sqrt), "invsqrt": (lambda a: np.abs(a) * 8, lambda x: 1 / np.sqrt(x)),
}

# the bit patterns (magnitudes) each function is also checked over: random
# FP32 values spread over its whole computing range, both signs where defined
SPAN = {"sin": 0x49800000, "cos": 0x49800000, "tan": 0x49800000, "asin": 0x3f800001,
        "acos": 0x3f800001, "atanh": 0x3f800000, "sinh": 0x42b30000, "cosh": 0x42b30000,
        "exp": 0x42d00000, "exp2": 0x43160000, "acosh": 0x7f800000, "log": 0x7f800000,
        "log2": 0x7f800000, "sqrt": 0x7f800000, "invsqrt": 0x7f800000}
POSITIVE = {"log", "log2", "sqrt", "invsqrt", "acosh"}


@pytest.mark.parametrize("fn", list(UNARY))
def test_unary_ulp_every_function(gpu, c1, fn):
    """Each of the 20 unary functions within 0.51 ulp of its FP64 value on
    the C1 operand mapped into the function's domain, and on 2^20 FP32 values
    drawn from its whole computing range (abs, sign, sqrt exactly)."""
    a, _ = c1
    mapf, ref = UNARY[fn]
    rng = np.random.default_rng(list(UNARY).index(fn))
    hi = SPAN.get(fn, 0x7f800000)
    bits = rng.integers(0x3f800000 if fn == "acosh" else 0, hi, 1 << 20, dtype=np.int64).astype(np.uint32)
    if fn not in POSITIVE:
        bits[::2] |= np.uint32(0x80000000)
    wide = bits.view(np.float32)
    for x in (mapf(a).astype(np.float32), wide):
        got = np.asarray(getattr(A(gpu, x), fn)())
        with np.errstate(all="ignore"):
            exact = ref(x.astype(np.float64))
        finite = np.isfinite(exact) & (np.abs(exact) <= np.finfo(np.float32).max)
        if fn in ("abs", "sign", "sqrt"):
            assert_bits(got, exact.astype(np.float32))
            continue
        err = ulp_err(got[finite], exact[finite])
        assert err.max() <= 0.51, (fn, float(err.max()), x[finite][np.argmax(err)])
        # overflow to inf and NaN where the FP64 value is out of range
        over = ~finite
        assert np.array_equal(np.isnan(got[over]), np.isnan(exact[over])), fn
        assert np.all(got[over & ~np.isnan(exact)] == exact[over & ~np.isnan(exact)].astype(np.float32)), fn


def test_c1_reductions(gpu, c1):
    a, _ = c1
    ga = A(gpu, a)
    s0 = np.asarray(ga.sum(axis=0))
    np.testing.assert_allclose(s0, a.astype(np.float64).sum(axis=0), rtol=0, atol=1e-4)
    np.testing.assert_allclose(s0, O.reduce_axis("sum", a, 0), rtol=0, atol=1e-4)
    m = np.asarray(ga.mean())
    assert abs(float(m[0]) - a.astype(np.float64).mean()) <= 1e-7
    assert_bits(ga.maximum(axis=1), a.max(axis=1))
    assert_bits(ga.minimum(axis=0), a.min(axis=0))


def test_c1_matmul(gpu):
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (512, 512)).astype(np.float32)
    y = rng.uniform(-1, 1, (512, 512)).astype(np.float32)
    got = np.asarray(A(gpu, x) @ A(gpu, y))
    ref = x.astype(np.float64) @ y.astype(np.float64)
    bound = np.abs(x).astype(np.float64) @ np.abs(y).astype(np.float64)
    assert (np.abs(got - ref) / bound).max() <= 512 * 2 ** -24


def test_c1_gather(gpu):
    rng = np.random.default_rng(0)
    src = rng.standard_normal(2 ** 20, dtype=np.float32)
    idx = rng.integers(0, 2 ** 20, 2 ** 18, dtype=np.uint32)
    got = A(gpu, src).gather(vk.U32Array(gpu, data=idx))
    assert_bits(got, src[idx])
    assert_bits(got, O.gather(src, idx))


# ---- broadcasting, random shapes ---------------------------------------------------
def _random_bcast_pair(rng):
    nd = int(rng.integers(1, 6))
    shape = [int(rng.integers(1, 7)) for _ in range(nd)]
    sa = [s if rng.random() < 0.6 else 1 for s in shape]
    sb = [s if rng.random() < 0.6 else 1 for s in shape]
    da = int(rng.integers(0, nd))
    db = int(rng.integers(0, nd))
    return tuple(sa[da:]) or (1,), tuple(sb[db:]) or (1,)


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div", "max", "min"])
def test_broadcast_random_shapes(gpu, op):
    rng = np.random.default_rng(123)
    fn = {"add": lambda x, y: x + y, "sub": lambda x, y: x - y, "mul": lambda x, y: x * y,
          "div": lambda x, y: x / y, "max": lambda x, y: x.max(y), "min": lambda x, y: x.min(y)}[op]
    for _ in range(40):
        sa, sb = _random_bcast_pair(rng)
        a = rng.standard_normal(sa).astype(np.float32)
        b = rng.standard_normal(sb).astype(np.float32)
        got = fn(A(gpu, a), A(gpu, b))
        assert_bits(got, O.binary_bcast(op, a, b))


def test_broadcast_c2_shapes_small(gpu):
    # the C2 operand patterns at a reduced size: (M,N)+(M,1), (M,N)+(1,N), in-place
    rng = np.random.default_rng(2)
    for M, N in ((64, 256), (33, 100), (1, 7), (5, 1)):
        a = rng.standard_normal((M, N)).astype(np.float32)
        col = rng.standard_normal((M, 1)).astype(np.float32)
        row = rng.standard_normal((1, N)).astype(np.float32)
        for op, f in (("add", np.add), ("sub", np.subtract), ("mul", np.multiply), ("div", np.divide)):
            assert_bits(A(gpu, a)._op(A(gpu, col), op), f(a, col))
            assert_bits(A(gpu, a)._op(A(gpu, row), op), f(a, row))
            assert_bits(A(gpu, col)._op(A(gpu, a), op), f(col, a))
            g = A(gpu, a)
            g._iop(A(gpu, row), op)
            assert_bits(g, f(a, row))
            g = A(gpu, a)
            g._iop(A(gpu, col), op)
            assert_bits(g, f(a, col))
            g = A(gpu, a)
            g._iop(1.25, op)
            assert_bits(g, f(a, np.float32(1.25)))


def test_broadcast_to_random(gpu):
    rng = np.random.default_rng(5)
    for _ in range(20):
        sa, _ = _random_bcast_pair(rng)
        target = tuple(int(rng.integers(2, 4)) if s == 1 else s for s in sa)
        lead = tuple(int(x) for x in rng.integers(1, 3, int(rng.integers(0, 3))))
        a = rng.standard_normal(sa).astype(np.float32)
        got = A(gpu, a).broadcast_to(lead + target)
        assert_bits(got, np.broadcast_to(a, lead + target))
        assert_bits(got, O.broadcast_to(a, lead + target))


def test_clamp_broadcast_random(gpu):
    rng = np.random.default_rng(6)
    for _ in range(20):
        sa, sb = _random_bcast_pair(rng)
        x = rng.standard_normal(sa).astype(np.float32)
        lo = (rng.standard_normal(sb) - 0.5).astype(np.float32)
        hi = (rng.standard_normal(sa) + 0.5).astype(np.float32)
        got = A(gpu, x).clamp(A(gpu, lo), A(gpu, hi))
        assert_bits(got, np.minimum(np.maximum(x, lo), hi))


# ---- reductions, random shapes ----------------------------------------------------
@pytest.mark.parametrize("op", ["sum", "prod", "maximum", "minimum", "mean"])
def test_reduce_axis_random(gpu, op):
    rng = np.random.default_rng(7)
    shapes = [(3, 5), (4, 70), (2, 3000), (5000, 3), (7, 6, 5), (2, 3, 4, 2), (130, 2, 300), (1, 9)]
    for shape in shapes:
        lo, hi = (0.99, 1.01) if op == "prod" else (0.5, 1.5)
        x = rng.uniform(lo, hi, shape).astype(np.float32)
        for axis in list(range(len(shape))) + [(0, len(shape) - 1)]:
            got = np.asarray(getattr(A(gpu, x), op)(axis=axis))
            if op in ("maximum", "minimum"):
                assert_bits(got, O.reduce_axis(op, x, axis))
                continue
            x64 = x.astype(np.float64)
            ref = {"sum": x64.sum, "prod": x64.prod, "mean": x64.mean}[op](axis=axis)
            k = x.size // max(np.asarray(ref).size, 1)
            np.testing.assert_allclose(got, ref, rtol=4 * k * 2 ** -24 + 1e-7)


def test_reduce_all_large_and_edge(gpu):
    rng = np.random.default_rng(8)
    for n in (1, 63, 64, 65, 4097, 100003, 1 << 22):
        x = rng.standard_normal(n).astype(np.float32)
        g = A(gpu, x)
        s = float(np.asarray(g.sum())[0])
        exact = float(np.sum(x.astype(np.float64)))
        assert abs(s - exact) <= abs(exact) * 2 ** -24 + 1e-6 * np.sqrt(n)
        assert float(np.asarray(g.maximum())[0]) == x.max()
        assert float(np.asarray(g.minimum())[0]) == x.min()
        mean = float(np.asarray(g.mean())[0])
        assert abs(mean - exact / n) <= abs(exact / n) * 2 ** -23 + 1e-12
    p = rng.uniform(0.999, 1.001, 1 << 20).astype(np.float32)
    got = float(np.asarray(A(gpu, p).prod())[0])
    ref = float(np.exp(np.sum(np.log(p.astype(np.float64)))))
    assert abs(got - ref) <= ref * 2 ** -23


def test_rebroadcast_random(gpu):
    rng = np.random.default_rng(9)
    for shape in ((4, 70), (2, 3000), (6, 5, 4), (300, 2, 3)):
        x = rng.uniform(0.5, 1.5, shape).astype(np.float32)
        for axis in range(len(shape)):
            got = A(gpu, x).maximum(axis=axis, rebroadcast=True)
            assert_bits(got, O.reduce_axis("maximum", x, axis, rebroadcast=True))
            got = np.asarray(A(gpu, x).sum(axis=axis, rebroadcast=True))
            ref = np.broadcast_to(x.astype(np.float64).sum(axis=axis, keepdims=True), shape)
            np.testing.assert_allclose(got, ref, rtol=shape[axis] * 2 ** -22)


def test_negative_axis_and_keepdims(gpu):
    x = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    g = A(gpu, x)
    np.testing.assert_allclose(g.sum(axis=-1), x.sum(axis=-1))
    np.testing.assert_allclose(g.sum(axis=(-1, 0), keepdims=True), x.sum(axis=(-1, 0), keepdims=True))
    assert g.sum(axis=(-1, 0), keepdims=True).shape == (1, 3, 1)


def test_setitem_getitem_and_flush(gpu):
    a = A(gpu, np.arange(6).reshape(2, 3))
    a[0, 1] = 10
    b = a + 1
    np.testing.assert_allclose(b, [[1, 11, 3], [4, 5, 6]])
    assert a[1, 2] == 5
    a[:] = 2.0
    np.testing.assert_allclose(a + 0, np.full((2, 3), 2.0))
    host = a.array
    host[1, 1] = -1
    a.flush()
    np.testing.assert_allclose(a * 1, [[2, 2, 2], [2, -1, 2]])
    z = vk.zeros(gpu, (3, 2))
    np.testing.assert_allclose(z, np.zeros((3, 2)))


def test_job_wait_semantics(gpu):
    a = A(gpu, np.ones(1 << 16))
    b = a + a
    assert b.job is not None
    b.job.wait(10 ** 10)
    b.wait()
    assert b.job is None
    gpu.wait()


def test_jobs_record_events_only_when_waited(gpu):
    """Jobs are tickets: enqueueing records no event; the first wait or query
    records one that covers every job issued so far, and later waits on
    covered jobs reuse it."""
    ctx = gpu._ctx
    gpu.wait()
    latest = ctx.latest
    x = A(gpu, np.arange(1 << 20, dtype=np.float32))
    results = [x * float(k) for k in range(1, 9)]
    assert ctx.latest is latest                       # no event per operation
    first, last = results[0].job, results[-1].job
    assert last._ticket > first._ticket
    first.wait()                                      # records one event ...
    e = ctx.latest
    assert e is not latest and e.ticket >= last._ticket
    assert last.done()                                # ... that covers later jobs
    last.wait()
    assert ctx.latest is e
    for k, r in enumerate(results, start=1):
        np.testing.assert_array_equal(np.asarray(r), np.arange(1 << 20, dtype=np.float32) * k)
    y = x + 1.0
    gpu.wait()                                        # a device wait completes every ticket
    assert y.job.done() and ctx.latest is e
    y.job.wait(1)


def test_timer_next_kernel(gpu):
    """vkh_timer_next_kernel: the next elementwise kernel records the pair at
    its own start and end.  Its result is unchanged; its time is positive and
    no longer than that of ordinary events recorded around the same launch;
    the pair is used once (a second op leaves it alone)."""
    import ctypes
    from vulkpy_amd import _lib
    rng = np.random.default_rng(31)
    a = rng.standard_normal((2048, 4096), dtype=np.float32)
    c = rng.standard_normal((2048, 1), dtype=np.float32)
    ga, gc = A(gpu, a), A(gpu, c)
    gpu.wait()
    h = gpu._handle
    s, e = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.call("vkh_timer_next_kernel", h, ctypes.byref(s), ctypes.byref(e))
    o0, o1 = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.call("vkh_timer_record", h, ctypes.byref(o0))
    r = ga + gc
    _lib.call("vkh_timer_record", h, ctypes.byref(o1))
    r2 = ga * 2.0  # not timed: the pair was used by the add
    ms, outer = ctypes.c_float(), ctypes.c_float()
    _lib.call("vkh_event_elapsed_ms", s, e, ctypes.byref(ms))
    _lib.call("vkh_event_elapsed_ms", o0, o1, ctypes.byref(outer))
    for ev in (s, e, o0, o1):
        _lib.lib.vkh_event_destroy(ev)
    np.testing.assert_array_equal(np.asarray(r), a + c)
    np.testing.assert_array_equal(np.asarray(r2), a * np.float32(2.0))
    assert 0.0 < ms.value <= outer.value, (ms.value, outer.value)
