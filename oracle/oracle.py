"""
Oracle for the vulkpy array / PRNG / nn path (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker or the timed CPU
baseline.  The product package ``vulkpy_amd`` never uses it.

Two parts:

* ``liboracle.so`` (``vk_oracle.c``): a C restatement of the reference's
  shaders and Xoshiro128++ host code, loaded with ctypes.
* NumPy restatements (below) of the nn formulas in the reference's FP32
  operation order.  Each function cites the reference file:line it follows
  (paths relative to the vulkpy source tree).

Pinning of each function (known answers, closed forms, invariants) lives in
``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_float, c_int, c_uint32, c_uint64, c_void_p, c_double
from typing import Iterable, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

BINOP = {"add": 0, "sub": 1, "mul": 2, "div": 3, "max": 4, "min": 5, "pow": 6,
         "rsub": 7, "rdiv": 8, "rpow": 9}
REDOP = {"sum": 0, "prod": 1, "maximum": 2, "minimum": 3}

_lib = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    P = c_void_p
    U64 = c_uint64
    sig = {
        "oracle_splitmix64": (c_uint64, [c_uint64]),
        "oracle_jump_as_written": (None, [P]),
        "oracle_jump_canonical": (None, [P]),
        "oracle_next": (c_uint32, [P]),
        "oracle_xoshiro_seed": (None, [c_uint32, U64, P]),
        "oracle_xoshiro_uint32": (None, [P, c_uint32, P, U64]),
        "oracle_xoshiro_float": (None, [P, c_uint32, P, U64]),
        "oracle_box_muller": (None, [P, P, U64, c_float, c_float]),
        "oracle_randrange": (None, [P, P, U64, c_uint32, c_uint32]),
        "oracle_binary": (None, [c_int, P, P, c_float, P, U64]),
        "oracle_binary_bcast": (None, [c_int, P, P, P, c_uint32, c_uint32, c_uint32, c_uint32, P]),
        "oracle_broadcast_to": (None, [P, P, c_uint32, c_uint32, c_uint32, P, P]),
        "oracle_reduce_axis": (None, [c_int, P, P, U64, U64, U64, c_int]),
        "oracle_matmul": (None, [P, P, P, U64, U64, U64]),
        "oracle_batch_affine": (None, [P, P, P, P, U64, U64, U64]),
        "oracle_gather": (None, [P, U64, P, P, U64]),
        "oracle_gather_axis": (None, [P, P, P, U64, U64, U64, U64]),
        "oracle_sum_f64": (c_double, [P, U64]),
        "oracle_num_threads": (c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


# ---- Xoshiro128++ (_vkarray.cc:577-719) ---------------------------------------
def splitmix64(x: int) -> int:
    """_vkarray.cc:589-595"""
    return int(load().oracle_splitmix64(x))


def jump_as_written(state: Sequence[int]) -> np.ndarray:
    """_vkarray.cc:597-621"""
    s = np.array(state, dtype=np.uint32)
    load().oracle_jump_as_written(_p(s))
    return s


def jump_canonical(state: Sequence[int]) -> np.ndarray:
    """Vigna's xoshiro128plusplus.c jump() (not what the reference does)"""
    s = np.array(state, dtype=np.uint32)
    load().oracle_jump_canonical(_p(s))
    return s


def next_state(state: np.ndarray) -> int:
    """one xoshiro128++ step (prng_xoshiro128pp_uint32.comp:32-42); advances state"""
    return int(load().oracle_next(_p(state)))


def xoshiro_seed(size: int, seed: int) -> np.ndarray:
    """[size, 4] uint32 initial states (_vkarray.cc:653-675)"""
    out = np.empty((size, 4), dtype=np.uint32)
    load().oracle_xoshiro_seed(size, seed, _p(out))
    return out


class Xoshiro128pp:
    """Reference generator restated on the CPU: chunked dispatch of
    _vkarray.cc:697-717 over the state buffer (shared by uint32 and float)."""

    def __init__(self, size: int = 64, seed: int = 0):
        self.size = size
        self.state = xoshiro_seed(size, seed)

    def randint(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.uint32)
        load().oracle_xoshiro_uint32(_p(self.state), self.size, _p(out), n)
        return out

    def random(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.float32)
        load().oracle_xoshiro_float(_p(self.state), self.size, _p(out), n)
        return out

    def normal(self, n: int, mean: float = 0.0, stddev: float = 1.0) -> np.ndarray:
        """random.py:92-124: n even -> n uniforms, n odd -> n + 1 uniforms"""
        u = self.random(n + (n % 2))
        return box_muller(u, n, mean, stddev)

    def randrange(self, n: int, low: int, high: int) -> np.ndarray:
        """random.py:157-186"""
        if low == 0 and high == 2 ** 32:
            return self.randint(n)
        u = self.random(n)
        return randrange(u, low, high - 1)


def box_muller(u: np.ndarray, n: int, mean: float, stddev: float) -> np.ndarray:
    """prng_box_muller.comp:19-32 over u (len >= n + n % 2)"""
    u = _f32(u)
    out = np.empty(n, dtype=np.float32)
    load().oracle_box_muller(_p(u), _p(out), n, mean, stddev)
    return out


def randrange(u: np.ndarray, low: int, high_incl: int) -> np.ndarray:
    """prng_randrange.comp:20-27"""
    u = _f32(u)
    out = np.empty(u.size, dtype=np.uint32)
    load().oracle_randrange(_p(u), _p(out), u.size, low, high_incl)
    return out


# ---- elementwise ----------------------------------------------------------------
def binary(op: str, a, b) -> np.ndarray:
    """same-shape (add.comp ...) or scalar (add_scalar.comp ...) op"""
    a = _f32(a)
    out = np.empty_like(a)
    if np.ndim(b) == 0 and not isinstance(b, np.ndarray):
        load().oracle_binary(BINOP[op], _p(a), None, float(b), _p(out), a.size)
    else:
        b = _f32(b)
        assert b.shape == a.shape
        load().oracle_binary(BINOP[op], _p(a), _p(b), 0.0, _p(out), a.size)
    return out


def binary_bcast(op: str, a, b) -> np.ndarray:
    """add_broadcast.comp:25-44 with the shape table of vkarray.py:499-516"""
    a, b = _f32(a), _f32(b)
    shape = np.broadcast_shapes(a.shape, b.shape)
    nd = len(shape)
    table = np.ones(3 * nd, dtype=np.uint32)
    table[nd - a.ndim:nd] = a.shape
    table[2 * nd - b.ndim:2 * nd] = b.shape
    table[2 * nd:] = shape
    out = np.empty(shape, dtype=np.float32)
    load().oracle_binary_bcast(BINOP[op], _p(a), _p(b), _p(out), a.size, b.size, out.size, nd, _p(table))
    return out


def broadcast_to(a, shape) -> np.ndarray:
    """broadcast.comp:25-44 with vkarray.py:1459-1472"""
    a = _f32(a)
    shape = tuple(int(s) for s in shape)
    a_shape = (1,) * (len(shape) - a.ndim) + a.shape
    sa = np.array(a_shape, dtype=np.uint32)
    sb = np.array(shape, dtype=np.uint32)
    out = np.empty(shape, dtype=np.float32)
    load().oracle_broadcast_to(_p(a), _p(out), a.size, out.size, len(shape), _p(sa), _p(sb))
    return out


# ---- reductions -------------------------------------------------------------------
def reduce_axis(op: str, a, axis, rebroadcast: bool = False) -> np.ndarray:
    """*_axis.comp / *_axis_rebroadcast.comp, one pass per axis, last axis first
    (vkarray.py:1194-1244), sequential FP32 over k"""
    a = _f32(a)
    if rebroadcast:
        ax = int(axis) % a.ndim
        prev = int(np.prod(a.shape[:ax], dtype=np.int64))
        post = int(np.prod(a.shape[ax + 1:], dtype=np.int64))
        out = np.empty_like(a)
        load().oracle_reduce_axis(REDOP[op], _p(a), _p(out), prev, a.shape[ax], post, 1)
        return out
    axes = sorted({int(x) % a.ndim for x in np.asarray(axis).reshape(-1)}, reverse=True)
    tmp = a
    for ax in axes:
        prev = int(np.prod(tmp.shape[:ax], dtype=np.int64))
        post = int(np.prod(tmp.shape[ax + 1:], dtype=np.int64))
        out = np.empty(tmp.shape[:ax] + tmp.shape[ax + 1:], dtype=np.float32)
        load().oracle_reduce_axis(REDOP[op], _p(tmp), _p(out), prev, tmp.shape[ax], post, 0)
        tmp = out
    return tmp


def mean_axis(a, axis) -> np.ndarray:
    """vkarray.py:1420-1431: sum over axis then `*= n_after / n_before` (FP32)"""
    a = _f32(a)
    s = reduce_axis("sum", a, axis)
    return (s * np.float32(s.size / a.size)).astype(np.float32)


# ---- matmul / gather --------------------------------------------------------------
def matmul(A, B) -> np.ndarray:
    """matmul.comp:23-33 (2-D)"""
    A, B = _f32(A), _f32(B)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    C = np.empty((M, N), dtype=np.float32)
    load().oracle_matmul(_p(A), _p(B), _p(C), M, K, N)
    return C


def batch_affine(W, b, x) -> np.ndarray:
    """batch_affine.comp:25-39"""
    W, b, x = _f32(W), _f32(b), _f32(x)
    out_dim, in_dim = W.shape
    B = x.shape[0]
    y = np.empty((B, out_dim), dtype=np.float32)
    load().oracle_batch_affine(_p(W), _p(b), _p(x), _p(y), B, in_dim, out_dim)
    return y


def gather(a, idx) -> np.ndarray:
    """gather.comp:21-26"""
    a = _f32(a).reshape(-1)
    idx = np.ascontiguousarray(np.asarray(idx, dtype=np.uint32))
    out = np.empty(idx.shape, dtype=np.float32)
    load().oracle_gather(_p(a), a.size, _p(idx), _p(out), idx.size)
    return out


def gather_axis(a, idx, axis: int) -> np.ndarray:
    """gather_axis.comp:24-43, output shape idx.shape + prev + post (vkarray.py:1506-1517)"""
    a = _f32(a)
    idx = np.ascontiguousarray(np.asarray(idx, dtype=np.uint32))
    axis = axis % a.ndim
    prev_shape, post_shape = a.shape[:axis], a.shape[axis + 1:]
    out = np.empty(idx.shape + prev_shape + post_shape, dtype=np.float32)
    load().oracle_gather_axis(_p(a), _p(idx), _p(out), int(np.prod(prev_shape, dtype=np.int64)),
                              a.shape[axis], int(np.prod(post_shape, dtype=np.int64)), idx.size)
    return out


def num_threads() -> int:
    return int(load().oracle_num_threads())


def sum_f64(a) -> float:
    a = _f32(a)
    return float(load().oracle_sum_f64(_p(a), a.size))


# ---- nn restatements (NumPy, FP32 operation order of the reference) ----------------
f32 = np.float32


def fexp(x):
    """FP64 exp rounded once to FP32 (the product's transcendental policy)"""
    return np.exp(np.asarray(x, dtype=np.float64)).astype(np.float32)


def flog(x):
    return np.log(np.asarray(x, dtype=np.float64)).astype(np.float32)


def np_max(a, b):
    """numpy.maximum with the ties-keep-first rule of the kernels"""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    return np.where((a >= b) | np.isnan(a), a, b).astype(np.float32)


def np_min(a, b):
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    return np.where((a <= b) | np.isnan(a), a, b).astype(np.float32)


def sign(x):
    x = np.asarray(x, dtype=np.float32)
    return np.where(x > 0, f32(1), np.where(x < 0, f32(-1), np.where(x == 0, f32(0), x))).astype(np.float32)


def relu(x):
    """nn/layers.py:185: x.max(0.0)"""
    return np_max(x, f32(0))


def relu_backward(y, dy):
    """nn/layers.py:207-210: max(sign(y), 0) * dy"""
    return (np_max(sign(y), f32(0)) * _f32(dy)).astype(np.float32)


def sigmoid(x):
    """nn/layers.py:240-244: 0 - x; exp; + 1; 1 / y"""
    y = (f32(0) - _f32(x)).astype(np.float32)
    y = fexp(y)
    y = (y + f32(1)).astype(np.float32)
    return (f32(1) / y).astype(np.float32)


def sigmoid_backward(y, dy):
    """nn/layers.py:264-267 (also Softmax.backward, 320-323)"""
    y = _f32(y)
    d = (f32(1) - y).astype(np.float32)
    d = (d * y).astype(np.float32)
    return (d * _f32(dy)).astype(np.float32)


def _seq_row_sum(v):
    """sequential FP32 sum over axis 1 (sum_axis.comp:28-30)"""
    v = _f32(v)
    acc = np.zeros(v.shape[0], dtype=np.float32)
    for j in range(v.shape[1]):
        acc = (acc + v[:, j]).astype(np.float32)
    return acc


def softmax(x):
    """nn/layers.py:297-300 with sequential row max / sum"""
    x = _f32(x)
    m = x[:, 0].copy()
    for j in range(x.shape[1]):
        m = np_max(m, x[:, j])
    e = fexp((x - m[:, None]).astype(np.float32))
    # denormal exponentials flushed to zero, as on the reference's Vulkan
    # drivers (test/test_nn.py:124-131 expects softmax([100, 0]) == [1, 0])
    e = np.where(e < np.finfo(np.float32).tiny, f32(0), e).astype(np.float32)
    s = _seq_row_sum(e)
    return (e / s[:, None]).astype(np.float32)


def cross_entropy_rows(x, y):
    """nn_cross_entropy.comp:26 + sum(axis=1): sum_j (-y) * log(x + 1e-8)"""
    x, y = _f32(x), _f32(y)
    t = ((-y) * flog((x + f32(1e-8)).astype(np.float32))).astype(np.float32)
    return _seq_row_sum(t)


def cross_entropy_backward(x, y, scale=1.0):
    """nn_cross_entropy_backward.comp:26 then `dx *= scale`"""
    x, y = _f32(x), _f32(y)
    d = ((-y) / (x + f32(1e-8)).astype(np.float32)).astype(np.float32)
    return (d * f32(scale)).astype(np.float32) if scale != 1.0 else d


def mse_rows(x, y):
    """nn/losses.py:294-296"""
    d = (_f32(y) - _f32(x)).astype(np.float32)
    return _seq_row_sum((d * d).astype(np.float32))


def huber_rows(x, y):
    """nn/losses.py:365-369"""
    d = np.abs((_f32(y) - _f32(x)).astype(np.float32))
    t = np_min(d, (d * d).astype(np.float32))
    t = (t * f32(0.5)).astype(np.float32)
    return _seq_row_sum(t)


def loss_grad(kind: str, x, y, scale=1.0):
    """nn/losses.py:249 (sce, x = softmax), 318-320 (mse), 391-393 (huber),
    then ReduceLoss.grad's `dx *= scale` (70-86)"""
    d = (_f32(x) - _f32(y)).astype(np.float32)
    if kind == "mse":
        d = (d * f32(2)).astype(np.float32)
    elif kind == "huber":
        d = np_min(np_max(d, f32(-1)), f32(1))
    if scale != 1.0:
        d = (d * f32(scale)).astype(np.float32)
    return d


def batch_mean(L):
    """ReduceLoss mean: L.mean(axis=0) = sum(axis=0) * f32(1/B) (vkarray.py:1429)"""
    L = _f32(L)
    s = f32(0)
    for v in L:
        s = f32(s + v)
    return f32(s * f32(1 / L.shape[0]))


def adam(w, g, m, v, lr, beta1, beta2, eps, b1t, b2t):
    """nn/optimizers.py:235-253 and nn/parameters.py:95 (b1t, b2t already
    multiplied by beta1, beta2 for this step); returns new (w, m, v)"""
    w, g, m, v = _f32(w), _f32(g), _f32(m), _f32(v)
    m = (m * f32(beta1)).astype(np.float32)
    m = (m + (g * f32(1 - beta1)).astype(np.float32)).astype(np.float32)
    v = (v * f32(beta2)).astype(np.float32)
    v = (v + ((g * g).astype(np.float32) * f32(1 - beta2)).astype(np.float32)).astype(np.float32)
    mhat = (m / f32(1 - b1t)).astype(np.float32)
    vhat = (v / f32(1 - b2t)).astype(np.float32)
    vhat = np.sqrt(vhat).astype(np.float32)
    vhat = (vhat + f32(eps)).astype(np.float32)
    mhat = (mhat * f32(-lr)).astype(np.float32)
    mhat = (mhat / vhat).astype(np.float32)
    return (w + mhat).astype(np.float32), m, v


def regularizer_grad(kind: str, w, l1: float = 0.0, l2: float = 0.0):
    """Regularizer.grad of nn/regularizers.py in FP32 per op: "lasso"
    coeff * sign(w) (56-69), "ridge" (2 coeff) * w with the scalar rounded
    once (113-127), "elastic" L1.grad + L2.grad (177-192)"""
    w = _f32(w)
    t1 = (sign(w) * f32(l1)).astype(np.float32)
    t2 = (w * f32(2 * l2)).astype(np.float32)
    if kind == "lasso":
        return t1
    if kind == "ridge":
        return t2
    return (t1 + t2).astype(np.float32)


def sgd(w, g, lr):
    """nn/optimizers.py:53"""
    return (_f32(w) + (_f32(g) * f32(-lr)).astype(np.float32)).astype(np.float32)


def adagrad(w, g, h, lr, eps):
    """nn/optimizers.py:134-140; returns (w, h)"""
    w, g, h = _f32(w), _f32(g), _f32(h)
    h = (h + (g * g).astype(np.float32)).astype(np.float32)
    s = np.sqrt(h).astype(np.float32)
    s = (s + f32(eps)).astype(np.float32)
    r = (g / s).astype(np.float32)
    r = (r * f32(-lr)).astype(np.float32)
    return (w + r).astype(np.float32), h


def dense_backward_reference(x, dy, W, chunk: int = 256):
    """nn/layers.py:126-141 restated: db = dy.sum(axis=0) (sequential),
    dW = sum_b dy[b, :, None] * x[b, None, :] (sequential over b, FP32),
    dx = dy @ W.  Processed in batch chunks instead of materialising the
    B x out x in product."""
    x, dy, W = _f32(x), _f32(dy), _f32(W)
    B, O = dy.shape
    I = x.shape[1]
    db = np.zeros(O, dtype=np.float32)
    dW = np.zeros((O, I), dtype=np.float32)
    for b in range(B):
        db = (db + dy[b]).astype(np.float32)
        dW = (dW + (dy[b][:, None] * x[b][None, :]).astype(np.float32)).astype(np.float32)
    dx = (dy.astype(np.float64) @ W.astype(np.float64)).astype(np.float32)
    return dW, db, dx
