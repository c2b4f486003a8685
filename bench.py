#!/usr/bin/env python3
"""
Bandwidth benchmark of the vulkpy array path on one MI355X.

Workload (one "step"): the 8192 x 16384 float32 suite -- A ~ N(0, 1) from
numpy.random.default_rng(0), broadcast operands col (8192, 1) and row
(1, 16384), a second full operand B, and P ~ U[0.999, 1.001] for products:

  * same-shape binary ops A (+ - * / max min) B, Q ** B        12 N bytes each
    (Q = |A| + 0.1, the positive base of the powers)
  * broadcast ops A (+ - * / max min) col and row, Q ** col,
    Q ** row                                                   8 N + |operand|
  * scalar ops A * 2, A + 1.5, Q ** 1.5, 1.5 ** A              8 N
  * in-place pairs: A += B, -= B; A += col, -= col;
    A *= row, /= row; A *= 2, /= 2; and Q **= 1                the ops' bytes
  * clamp(A, -0.5, 0.5) and the 20 unary functions on their
    domains (A; D ~ U[-1, 1) for asin / acos / atanh; Q for
    log / log2 / sqrt / invsqrt; E = 1 + |A| for acosh)        8 N each
  * sum / maximum / minimum / mean over axis 0, 1 and None,
    prod of P over axis 0, 1 and None                         4 N + |output|
  * Xoshiro128++ with 2^20 states: 2^28 uint32, float, normal 4 * 2^28 + 32 * 2^20

After the timed steps every op's result is compared once with FP64 NumPy on
a sample (its first 64 rows; every output of axis-0 and full reductions; the
first 2^22 outputs of fresh generators against the oracle):
"per_op_max_ulp" is the largest error in units of the FP32 spacing at the
FP64 value (0 for the generators' bit-exact uint32 and float streams); for
sums and means the unit is 2^-24 times the sum of the magnitudes reduced
(divided by their count for means), for products 2^-24 times the value, so
a length-k FP32 recurrence is bounded by k.

value = algorithmic bytes of all ops / wall time of the timed steps (GB/s,
inputs resident in HBM).  The dominant kernel (broadcast A + col) is also
timed per launch in the timed region for the roofline, with HIP events that
its own dispatch records at the kernel's start and end on the library's
stream (vkh_timer_next_kernel).
The CPU baseline times the oracle (C + OpenMP, the reference's algorithms) on
the first 512 rows of the same operands, and NumPy beside it.

Two more sections of the same JSON line cover the rest of the metric:
  * "mlp": the 256-4096-4096-1 regression MLP (batch 16384, Huber, Ridge,
    Adam), plain FP32 step against the BF16 step captured as one hipGraph:
    steps/s of each, the speed-up and the final losses after the same number
    of steps, and the matrix-core utilisation of each product of the step;
    its "seed_study" repeats the FP32 and BF16 runs over R weight and data
    seeds (plus BF16 arms with one product family in split-BF16) and reports
    the losses' means and standard deviations;
  * "gemm": 8192^3 and 65536 x 1024 x 1024 products in FP32, BF16 and
    split-BF16, TFLOP/s and error against an FP64 product.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                [--mlp-steps S] [--mlp-seeds R] [--no-gemm] [--no-c2]
                [--pmc-kernels | --gemm-pmc [fp32|bf16|bf16x3]]   (counter runs only)
                [--mlp-trace S]   (kernel traces of the MLP steps only)
                [--mlp-trace-graphs S]   (the same as graph replays)
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "HBM GB/s (elementwise/reduce); MLP train steps/s, BF16-MFMA vs FP32 at tol"
M_ROWS, N_COLS = 8192, 16384
N = M_ROWS * N_COLS
PRNG_STATES, PRNG_N = 1 << 20, 1 << 28
PEAK_HBM_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


class Timer:
    """HIP timing events around one launch on the library's stream."""

    def __init__(self, gpu):
        from vulkpy_amd import _lib
        self._lib = _lib
        self.gpu = gpu
        self.pairs = []

    def start(self):
        ev = ctypes.c_void_p()
        self._lib.call("vkh_timer_record", self.gpu._handle, ctypes.byref(ev))
        return ev.value

    def stop(self, start):
        ev = ctypes.c_void_p()
        self._lib.call("vkh_timer_record", self.gpu._handle, ctypes.byref(ev))
        self.pairs.append((start, ev.value))

    def next_kernel(self):
        """Time the next elementwise kernel alone: its own dispatch records
        the pair at the kernel's start and end (no event records around it)."""
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        self._lib.call("vkh_timer_next_kernel", self.gpu._handle, ctypes.byref(a), ctypes.byref(b))
        self.pairs.append((a.value, b.value))

    def mean_ms(self) -> float:
        ms = ctypes.c_float()
        total = 0.0
        for a, b in self.pairs:
            self._lib.call("vkh_event_elapsed_ms", a, b, ctypes.byref(ms))
            total += ms.value
            self._lib.lib.vkh_event_destroy(a)
            self._lib.lib.vkh_event_destroy(b)
        n = len(self.pairs)
        self.pairs = []
        return total / max(n, 1)


S_ROWS = 64  # rows of the 8192 x 16384 operands in the per-op error sample

# unary functions of the suite: (input name, FP64 reference)
UNARY_SUITE = {
    "abs": ("A", np.abs), "sign": ("A", np.sign), "sin": ("A", np.sin), "cos": ("A", np.cos),
    "tan": ("A", np.tan), "asin": ("D", np.arcsin), "acos": ("D", np.arccos),
    "atan": ("A", np.arctan), "sinh": ("A", np.sinh), "cosh": ("A", np.cosh),
    "tanh": ("A", np.tanh), "asinh": ("A", np.arcsinh), "acosh": ("E", np.arccosh),
    "atanh": ("D", np.arctanh), "exp": ("A", np.exp), "log": ("Q", np.log),
    "exp2": ("A", np.exp2), "log2": ("Q", np.log2), "sqrt": ("Q", np.sqrt),
    "invsqrt": ("Q", lambda x: 1.0 / np.sqrt(x)),
}

F64_BIN = {"add": np.add, "sub": np.subtract, "mul": np.multiply, "div": np.divide,
           "max": np.maximum, "min": np.minimum, "pow": np.power}


def build_suite(vk, gpu, ops_in, rng_gen, ubuf, fbuf, timer):
    """
    The operations of one step: a list of (name, algorithmic bytes, run,
    check).  ``run()`` enqueues the op; ``check()`` (used once, after the
    timed steps) runs it again without changing any suite operand and returns
    (result sample, FP64 value of the same elements): the first S_ROWS rows
    for elementwise ops and axis-1 reductions, every output of axis-0 and
    full reductions.
    """
    from vulkpy_amd import _lib
    A, B, col, row, P = (ops_in[k] for k in ("A", "B", "col", "row", "P"))
    Q, D, E = (ops_in[k] for k in ("Q", "D", "E"))
    ops = []
    colb, rowb = 4 * M_ROWS, 4 * N_COLS

    def head(arr, n=None):
        """first n elements (default: S_ROWS rows) of a device array, as FP64"""
        n = n if n is not None else min(arr.size, S_ROWS * N_COLS)
        out = np.empty(n, dtype=np.float32)
        _lib.call("vkh_download", gpu._handle, out.ctypes.data, arr._ptr, out.nbytes)
        return out.astype(np.float64)

    def rows(arr):
        return head(arr).reshape(S_ROWS, N_COLS)

    def colv():
        return head(col, S_ROWS).reshape(S_ROWS, 1)

    def rowv():
        return head(row, N_COLS).reshape(1, N_COLS)

    def add(name, nbytes, fn, check=None):
        ops.append((name, nbytes, fn, check))

    binf = {"add": lambda x, y: x + y, "sub": lambda x, y: x - y, "mul": lambda x, y: x * y,
            "div": lambda x, y: x / y, "max": lambda x, y: x.max(y), "min": lambda x, y: x.min(y),
            "pow": lambda x, y: x ** y}
    # same-shape binary ops (pow on the positive base Q = |A| + 0.1)
    for o, f in binf.items():
        X = Q if o == "pow" else A
        add(f"{o}_same", 12 * N, (lambda f, X: lambda: f(X, B))(f, X),
            (lambda f, X, o: lambda: (rows(f(X, B)), F64_BIN[o](rows(X), rows(B))))(f, X, o))
    # broadcast binary ops against the (8192, 1) column and the (1, 16384) row
    for o, f in binf.items():
        X = Q if o == "pow" else A
        if o == "add":
            def timed(f=f):
                timer.next_kernel()
                return f(A, col)
            add("add_col", 8 * N + colb, timed, lambda: (rows(A + col), rows(A) + colv()))
        else:
            add(f"{o}_col", 8 * N + colb, (lambda f, X: lambda: f(X, col))(f, X),
                (lambda f, X, o: lambda: (rows(f(X, col)), F64_BIN[o](rows(X), colv())))(f, X, o))
        add(f"{o}_row", 8 * N + rowb, (lambda f, X: lambda: f(X, row))(f, X),
            (lambda f, X, o: lambda: (rows(f(X, row)), F64_BIN[o](rows(X), rowv())))(f, X, o))
    # scalar operands
    add("mul_scalar", 8 * N, lambda: A * 2.0, lambda: (rows(A * 2.0), rows(A) * 2.0))
    add("add_scalar", 8 * N, lambda: A + 1.5, lambda: (rows(A + 1.5), rows(A) + 1.5))
    add("pow_scalar", 8 * N, lambda: Q ** 1.5, lambda: (rows(Q ** 1.5), rows(Q) ** 1.5))
    add("rpow_scalar", 8 * N, lambda: 1.5 ** A, lambda: (rows(1.5 ** A), 1.5 ** rows(A)))

    # in-place ops, in pairs that leave the operands (nearly) as they were;
    # each check applies the pair's first op to a copy
    def inplace_check(X, op, other, ref_other):
        def check():
            Z = X * 1.0
            getattr(Z, op)(other)
            return rows(Z), F64_BIN[{"__iadd__": "add", "__isub__": "sub", "__imul__": "mul",
                                     "__itruediv__": "div", "__ipow__": "pow"}[op]](rows(X), ref_other())
        return check

    def pair(X, op1, op2, other):
        def run():
            getattr(X, op1)(other)
            getattr(X, op2)(other)
        return run

    add("inplace_same_x2", 2 * 12 * N, pair(A, "__iadd__", "__isub__", B),
        inplace_check(A, "__iadd__", B, lambda: rows(B)))
    add("inplace_col_x2", 2 * (8 * N + colb), pair(A, "__iadd__", "__isub__", col),
        inplace_check(A, "__iadd__", col, colv))
    add("inplace_row_x2", 2 * (8 * N + rowb), pair(A, "__imul__", "__itruediv__", row),
        inplace_check(A, "__imul__", row, rowv))
    add("inplace_scalar_x2", 2 * 8 * N, pair(A, "__imul__", "__itruediv__", 2.0),
        inplace_check(A, "__imul__", 2.0, lambda: 2.0))
    add("inplace_pow_scalar", 8 * N, lambda: Q.__ipow__(1.0),
        inplace_check(Q, "__ipow__", 1.0, lambda: 1.0))
    # clamp with scalar bounds
    add("clamp", 8 * N, lambda: A.clamp(-0.5, 0.5),
        lambda: (rows(A.clamp(-0.5, 0.5)), np.clip(rows(A), -0.5, 0.5)))
    # unary math
    for fn, (src, ref) in UNARY_SUITE.items():
        X = ops_in[src]
        add(fn, 8 * N, (lambda fn, X: lambda: getattr(X, fn)())(fn, X),
            (lambda fn, X, ref: lambda: (rows(getattr(X, fn)()), ref(rows(X))))(fn, X, ref))

    # reductions (prod on P ~ U[0.999, 1.001])
    full_cache = {}

    def full_f64(X):
        """FP64 column sums, sums of magnitudes, log-products, maxima and
        minima of a whole operand (read in row blocks; kept for the operand's
        other reductions)"""
        if id(X) in full_cache:
            return full_cache[id(X)]
        cs, ca, pr = np.zeros(N_COLS), np.zeros(N_COLS), np.zeros(N_COLS)
        mx, mn = np.full(N_COLS, -np.inf), np.full(N_COLS, np.inf)
        step_rows = 512
        for r0 in range(0, M_ROWS, step_rows):
            blk = head_rows(X, r0, step_rows)
            cs += blk.sum(axis=0)
            ca += np.abs(blk).sum(axis=0)
            pr += np.log(np.abs(blk)).sum(axis=0)
            mx = np.maximum(mx, blk.max(axis=0))
            mn = np.minimum(mn, blk.min(axis=0))
        full_cache[id(X)] = (cs, ca, pr, mx, mn)
        return full_cache[id(X)]

    def head_rows(X, r0, nr):
        out = np.empty(nr * N_COLS, dtype=np.float32)
        _lib.call("vkh_download", gpu._handle, out.ctypes.data, X._ptr + 4 * r0 * N_COLS, out.nbytes)
        return out.astype(np.float64).reshape(nr, N_COLS)

    def red_check(X, f, axis):
        """(result, FP64 value, scale): sums and means are measured against
        2^-24 times the sum of magnitudes of the reduced elements (mean: over
        their count), products against 2^-24 times the value, so a length-k
        FP32 recurrence stays within k; maxima and minima in ulps"""
        def check():
            got = np.asarray(getattr(X, f)(axis=axis), dtype=np.float64).reshape(-1)
            if axis == 1:
                r = rows(X)
                exact = {"sum": r.sum(1), "maximum": r.max(1), "minimum": r.min(1), "mean": r.mean(1),
                         "prod": np.exp(np.log(r).sum(1))}[f]
                mag = {"sum": np.abs(r).sum(1), "mean": np.abs(r).mean(1), "prod": np.abs(exact)}.get(f)
                return got[:S_ROWS], exact, mag
            cs, ca, lp, mx, mn = full_f64(X)
            if axis == 0:
                exact = {"sum": cs, "maximum": mx, "minimum": mn, "mean": cs / M_ROWS, "prod": np.exp(lp)}[f]
                mag = {"sum": ca, "mean": ca / M_ROWS, "prod": np.abs(exact)}.get(f)
                return got, exact, mag
            exact = np.array([{"sum": cs.sum(), "maximum": mx.max(), "minimum": mn.min(), "mean": cs.sum() / N,
                               "prod": np.exp(lp.sum())}[f]])
            mag = {"sum": np.array([ca.sum()]), "mean": np.array([ca.sum() / N]), "prod": np.abs(exact)}.get(f)
            return got, exact, mag
        return check

    for f in ("sum", "maximum", "minimum", "mean", "prod"):
        X = P if f == "prod" else A
        add(f"{f}_axis0", 4 * N + 4 * N_COLS, (lambda f, X: lambda: getattr(X, f)(axis=0))(f, X),
            red_check(X, f, 0))
        add(f"{f}_axis1", 4 * N + 4 * M_ROWS, (lambda f, X: lambda: getattr(X, f)(axis=1))(f, X),
            red_check(X, f, 1))
        add(f"{f}_all", 4 * N + 4, (lambda f, X: lambda: getattr(X, f)())(f, X), red_check(X, f, None))

    # Xoshiro128++ with 2^20 states: checked on fresh generators of the same
    # seed against the oracle's first 2^22 outputs (see prng_check)
    prng_bytes = 4 * PRNG_N + 32 * PRNG_STATES
    add("xoshiro_uint32", prng_bytes, lambda: rng_gen.randint(buffer=ubuf), "prng")
    add("xoshiro_float", prng_bytes, lambda: rng_gen.random(buffer=fbuf), "prng")
    add("xoshiro_normal", prng_bytes, lambda: rng_gen.normal(buffer=fbuf), "prng")
    return ops


def max_ulp(got, exact) -> float:
    """largest |got - exact| in units of the FP32 spacing at |exact|"""
    got = np.asarray(got, dtype=np.float64)
    exact = np.asarray(exact, dtype=np.float64)
    ok = np.isfinite(exact)
    if not ok.all():  # inf / NaN must match exactly
        bad = ~ok & ~((got == exact) | (np.isnan(got) & np.isnan(exact)))
        if bad.any():
            return float("inf")
    sp = np.spacing(np.abs(exact[ok]).astype(np.float32)).astype(np.float64)
    if not ok.any():
        return 0.0
    return float((np.abs(got[ok] - exact[ok]) / sp).max())


def prng_check(vk, gpu, n_check: int = 1 << 22) -> dict:
    """Max ulp (0 = bit for bit) of the three generator ops against the oracle:
    each on a fresh Xoshiro128pp(size=2^20, seed=1) drawing the suite's 2^28
    outputs, whose first n_check (four per state) the oracle restates."""
    from oracle import oracle as O
    out = {}
    for name in ("xoshiro_uint32", "xoshiro_float", "xoshiro_normal"):
        g = vk.random.Xoshiro128pp(gpu, size=PRNG_STATES, seed=1)
        o = O.Xoshiro128pp(PRNG_STATES, 1)
        if name == "xoshiro_uint32":
            got = np.asarray(g.randint(shape=(PRNG_N,)))[:n_check]
            out[name] = 0.0 if np.array_equal(got, o.randint(n_check)) else float("inf")
        elif name == "xoshiro_float":
            got = np.asarray(g.random(shape=(PRNG_N,)))[:n_check]
            want = o.random(n_check)
            out[name] = 0.0 if np.array_equal(got.view(np.uint32), want.view(np.uint32)) else float("inf")
        else:
            got = np.asarray(g.normal(shape=(PRNG_N,)))[:n_check]
            out[name] = round(max_ulp(got, o.normal(n_check).astype(np.float64)), 3)
        del g
        gpu.wait()
    return out


PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0, "bf16x3": 2500.0 / 3}  # split: 3 BF16 products


def mlp_gemm_utilisation(vk, gpu, timer, B: int, dims, reps: int = 3) -> dict:
    """
    Matrix-core utilisation of each product of the MLP step: the forward,
    weight-gradient and input-gradient products of every Dense layer (the
    first layer's input gradient is not computed), timed alone with HIP events
    over `reps` launches behind an untimed one, in the configuration the step
    uses (configurations chosen by autotuning stay on the context), operand
    packing included.  frac = TFLOP/s / the mode's dense peak (split-BF16: a
    third of the BF16 peak, since it runs three products).  Products with
    min(M, N) <= 16 or K <= 16 run on the vector units (path "vector").
    """
    from vulkpy_amd import _lib
    rng = np.random.default_rng(7)
    shapes = []
    for i, (din, dout) in enumerate(dims):
        shapes.append((f"forward {i + 1}", 0, 1, B, dout, din))
        shapes.append((f"weight gradient {i + 1}", 1, 0, dout, din, B))
        if i > 0:
            shapes.append((f"input gradient {i + 1}", 0, 0, B, din, dout))
    out = {}
    for mode in ("fp32", "bf16", "bf16x3"):
        rows = []
        for name, ta, tb, M, N, K in shapes:
            A = vk.Array(gpu, data=rng.standard_normal((K, M) if ta else (M, K), dtype=np.float32))
            Bm = vk.Array(gpu, data=rng.standard_normal((N, K) if tb else (K, N), dtype=np.float32))
            C = vk.Array(gpu, shape=(M, N))

            def run():
                _lib.call("vkh_gemm", gpu._handle, _lib.GEMM_MODE[mode], ta, tb, M, N, K, 1.0,
                          A._ptr, M if ta else K, Bm._ptr, K if tb else N, 0.0, C._ptr, N,
                          None, 0, None)
            run()
            gpu.wait()
            run()
            t = timer.start()
            for _ in range(reps):
                run()
            timer.stop(t)
            ms = timer.mean_ms() / reps
            tflops = 2.0 * M * N * K / (ms * 1e-3) / 1e12
            rows.append({"product": name, "M": M, "N": N, "K": K,
                         "path": "vector" if min(M, N) <= 16 or K <= 16 else "mfma",
                         "ms": round(ms, 4), "tflops": round(tflops, 1),
                         "frac": round(tflops / PEAK_TFLOPS[mode], 3)})
            del A, Bm, C
        out[mode] = rows
        gpu.wait()
    return out


def mlp_compare(vk, gpu, steps: int, warmup: int = 2, uncaptured: bool = False, timer=None,
                graphs_only: bool = False):
    """
    Regression MLP Dense(256->4096)-ReLU-Dense(4096->4096)-ReLU-Dense(4096->1),
    HuberLoss(delta=1), Ridge(1e-4) on every weight (applied, see
    ``regularize``), Adam lr=1e-3, batch 16384, x ~ N(0, 1),
    y = sin(x @ P) + 0.1 N(0, 1) with a fixed projection P (256 x 1).
    Every arm starts from the same HeNormal weights and trains `steps` steps
    on the same batch: the plain FP32 step (nn.Sequence, one launch per op,
    FP32 matrix cores) and the step captured as one hipGraph
    (nn.GraphSequence) with BF16 and with split-BF16 (3-product) matrix-core
    GEMMs, FP32 master weights and Adam in both, and kernel configurations
    autotuned per product shape before the capture.  steps/s is timed over
    the steps after the first `warmup` (the graph arms tune and capture on
    their first step).  Besides the final training loss, each arm's trained
    weights are evaluated in FP32 on a held-out batch of the same
    distribution.  With `uncaptured` only the BF16 and split-BF16 steps run,
    their kernels launched one at a time instead of as a graph replay; with
    `graphs_only` only those two steps run, as graph replays.
    """
    from vulkpy_amd import nn
    B, D, H = 16384, 256, 4096
    rng = np.random.default_rng(5)
    x = rng.standard_normal((B, D), dtype=np.float32)
    proj = (rng.standard_normal((D, 1)) / np.sqrt(D)).astype(np.float32)
    y = (np.sin(x @ proj) + 0.1 * rng.standard_normal((B, 1))).astype(np.float32)
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    xt = rng.standard_normal((B, D), dtype=np.float32)
    yt = (np.sin(xt @ proj) + 0.1 * rng.standard_normal((B, 1))).astype(np.float32)
    Xt, Yt = vk.Array(gpu, data=xt), vk.Array(gpu, data=yt)
    dims = [(D, H), (H, H), (H, 1)]

    def layers(precision):
        out = []
        for k, (i, o) in enumerate(dims):
            out.append(nn.Dense(gpu, i, o, w_init=nn.HeNormal(gpu, i, seed=100 + k),
                                w_opt=nn.Adam(gpu, lr=1e-3), b_opt=nn.Adam(gpu, lr=1e-3),
                                w_reg=nn.Ridge(1e-4), precision=precision))
            if k + 1 < len(dims):
                out.append(nn.ReLU())
        return out

    checkpoints = sorted({1, 10, 50, 100, 200, 300, 400, steps} & set(range(1, steps + 1)))

    def run(model):
        losses = {}
        t0 = None
        for s in range(steps):
            if s == warmup:
                gpu.wait()
                t0 = time.perf_counter()
            _, L = model.train(X, Y)
            if s + 1 in checkpoints:  # a handful of scalar reads
                losses[s + 1] = float(np.asarray(L))
        gpu.wait()
        dt = time.perf_counter() - t0
        return (steps - warmup) / dt, losses

    def heldout(model):
        """loss of the trained weights on (Xt, Yt), FP32 products"""
        for L in model.L:
            if isinstance(L, nn.Dense):
                L.precision = "fp32"
        h = Xt
        for L in model.L:
            h = L(h)
        return float(np.asarray(nn.HuberLoss()(h, Yt)))

    # The control arm repeats the FP32 step with one change of summation
    # order only: the bias-gradient column sums of the 4096-wide layers split
    # the batch into 256 chunks instead of 128 (a reduction configuration).
    # How far its losses move from the FP32 arm's is the noise floor that
    # any change of arithmetic order meets on this run.
    from vulkpy_amd import _lib
    bias_shape = (1, 1, B, H)
    reorder_cfg = next(i for i, c in enumerate(vk.reduce_configs("cols")) if c["par"] == 1024)

    def reorder(on: bool):
        _lib.call("vkh_reduce_set_config", gpu._handle, *bias_shape, reorder_cfg if on else -1)

    arms = (("fp32", lambda: nn.Sequence(layers("fp32"), nn.HuberLoss(), regularize=True), None),
            ("fp32_reordered", lambda: nn.Sequence(layers("fp32"), nn.HuberLoss(), regularize=True),
             reorder),
            ("bf16_graph", lambda: nn.GraphSequence(layers("bf16"), nn.HuberLoss(),
                                                    precision="bf16", regularize=True,
                                                    autotune=True), None),
            ("bf16x3_graph", lambda: nn.GraphSequence(layers("bf16x3"), nn.HuberLoss(),
                                                      precision="bf16x3", regularize=True,
                                                      autotune=True), None))
    if graphs_only:
        arms = arms[2:]
    elif uncaptured:
        # the graph arms' kernels launched one at a time (for kernel traces)
        arms = tuple((name.replace("graph", "uncaptured"),
                      lambda p=name.split("_")[0]: nn.GraphSequence(
                          layers(p), nn.HuberLoss(), precision=p, regularize=True,
                          autotune=True, capture=False), None)
                     for name, _, _ in arms[2:])
    res = {}
    for name, model, hook in arms:
        log(f"bench: MLP {name} {steps} steps")
        if hook:
            hook(True)
        m = model()
        sps, losses = run(m)
        res[name] = {"steps_per_s": round(sps, 2), "loss_first": losses[1], "loss_final": losses[steps],
                     "loss_at_step": {str(k): v for k, v in losses.items()},
                     "heldout_loss": heldout(m)}
        if hook:
            hook(False)
        log(f"bench: MLP {name} {res[name]}")
        del m
        gpu.wait()
    out = {
        "workload": ("C5 regression MLP 256-4096-4096-1 ReLU, HuberLoss(delta=1), Ridge(1e-4) on "
                     "weights, Adam lr=1e-3, batch 16384, x~N(0,1), y=sin(x@P)+0.1N(0,1)"),
        "steps": steps,
        "timed_steps": steps - warmup,
        "gemm_flop_per_step": 2 * (B * H * D + B * H * H + B * H) * 2 + 2 * (B * H * H + B * H),
    }
    if uncaptured or graphs_only:
        out.update(res)
        return out
    if timer is not None:
        log("bench: MLP per-product matrix-core utilisation")
        out["gemm_utilisation"] = mlp_gemm_utilisation(vk, gpu, timer, B, dims)
    ref = res["fp32"]
    for name, r in res.items():
        out[name] = r
        if name == "fp32":
            continue
        out[name]["speedup"] = round(r["steps_per_s"] / ref["steps_per_s"], 3)
        out[name]["final_loss_rel_diff"] = round(abs(r["loss_final"] - ref["loss_final"]) / abs(ref["loss_final"]), 5)
        out[name]["heldout_rel_diff"] = round(abs(r["heldout_loss"] - ref["heldout_loss"]) / abs(ref["heldout_loss"]), 5)
        out[name]["rel_diff_at_step"] = {
            k: round(abs(v - ref["loss_at_step"][k]) / abs(ref["loss_at_step"][k]), 5)
            for k, v in r["loss_at_step"].items()}
        out[name]["meets_final_loss_tol"] = out[name]["final_loss_rel_diff"] <= 0.01
    return out


SEED_ARMS = (
    ("fp32", None),
    ("bf16", "bf16"),
    ("bf16_wgrad_split", {"forward": "bf16", "input_grad": "bf16", "weight_grad": "bf16x3"}),
    ("bf16_dgrad_split", {"forward": "bf16", "input_grad": "bf16x3", "weight_grad": "bf16"}),
    ("bf16_fwd_split", {"forward": "bf16x3", "input_grad": "bf16", "weight_grad": "bf16"}),
)


def mlp_seed_study(vk, gpu, seeds, steps: int = 500) -> dict:
    """
    The regression MLP of mlp_compare over several weight and data seeds:
    for each seed, fresh inputs, targets, held-out batch and HeNormal weights,
    trained `steps` steps by the plain FP32 step (nn.Sequence) and by captured
    steps (nn.GraphSequence) whose products are BF16, or BF16 with one product
    family (weight gradients, input gradients or forward) in split-BF16.
    Reports each arm's final training loss, held-out loss and loss at step
    100 per seed, their mean and standard deviation over the seeds, and
    whether each BF16 arm's mean lies within one FP32 standard deviation of
    the FP32 mean and within 1% of it.
    """
    from vulkpy_amd import nn
    B, D, H = 16384, 256, 4096
    dims = [(D, H), (H, H), (H, 1)]
    per = {name: {"final": [], "heldout": [], "step100": []} for name, _ in SEED_ARMS}
    for seed in seeds:
        rng = np.random.default_rng(1000 + seed)
        x = rng.standard_normal((B, D), dtype=np.float32)
        proj = (rng.standard_normal((D, 1)) / np.sqrt(D)).astype(np.float32)
        y = (np.sin(x @ proj) + 0.1 * rng.standard_normal((B, 1))).astype(np.float32)
        xt = rng.standard_normal((B, D), dtype=np.float32)
        yt = (np.sin(xt @ proj) + 0.1 * rng.standard_normal((B, 1))).astype(np.float32)
        X, Y, Xt, Yt = (vk.Array(gpu, data=v) for v in (x, y, xt, yt))

        def layers(precision):
            out = []
            for k, (i, o) in enumerate(dims):
                out.append(nn.Dense(gpu, i, o, w_init=nn.HeNormal(gpu, i, seed=10 * seed + k),
                                    w_opt=nn.Adam(gpu, lr=1e-3), b_opt=nn.Adam(gpu, lr=1e-3),
                                    w_reg=nn.Ridge(1e-4), precision=precision))
                if k + 1 < len(dims):
                    out.append(nn.ReLU())
            return out

        for name, prec in SEED_ARMS:
            if prec is None:
                model = nn.Sequence(layers("fp32"), nn.HuberLoss(), regularize=True)
            else:
                model = nn.GraphSequence(layers("bf16"), nn.HuberLoss(), precision=prec,
                                         regularize=True, autotune=True)
            l100 = None
            for s in range(steps):
                _, L = model.train(X, Y)
                if s + 1 == 100:
                    l100 = float(np.asarray(L))
            final = float(np.asarray(L))
            for Lr in model.L:
                if isinstance(Lr, nn.Dense):
                    Lr.precision = "fp32"
            h = Xt
            for Lr in model.L:
                h = Lr(h)
            held = float(np.asarray(nn.HuberLoss()(h, Yt)))
            per[name]["final"].append(final)
            per[name]["heldout"].append(held)
            per[name]["step100"].append(l100)
            log(f"bench: MLP seed {seed} {name}: final {final:.6g} heldout {held:.6g} step100 {l100:.6g}")
            del model
            gpu.wait()
    out = {"seeds": list(seeds), "steps": steps}
    ref = {k: (float(np.mean(v)), float(np.std(v, ddof=1))) for k, v in per["fp32"].items()}
    for name, _ in SEED_ARMS:
        arm = {}
        for k, v in per[name].items():
            mean, std = float(np.mean(v)), float(np.std(v, ddof=1))
            arm[k] = {"per_seed": [round(t, 8) for t in v], "mean": round(mean, 8), "std": round(std, 8)}
            if name != "fp32":
                rm, rs = ref[k]
                arm[k]["mean_rel_diff"] = round(abs(mean - rm) / rm, 5)
                arm[k]["within_fp32_std"] = abs(mean - rm) <= rs
                arm[k]["within_1pct"] = abs(mean - rm) <= 0.01 * rm
        out[name] = arm
    return out


def gemm_report(vk, gpu, timer, reps: int = 3, sample_every: int = 128):
    """
    C3 GEMMs: C = A @ B for 8192^3 and 65536 x 1024 x 1024 in each mode, with
    N(0, 1) inputs and with U[-1, 1] inputs whose every 64th row of A is
    scaled by 1e3.  TFLOP/s = 2MNK / mean launch time (HIP events; operand
    packing included).  Errors against an FP64 product on every
    `sample_every`-th row: max and mean of |C - C64| / (|A| |B|) elementwise,
    and ||C - C64||_F / ||C64||_F.
    """
    out = []
    timer.mean_ms()  # start from an empty list of event pairs
    rng = np.random.default_rng(3)
    for (M, Nn, K) in ((8192, 8192, 8192), (65536, 1024, 1024)):
        for dist in ("normal", "range"):
            if dist == "normal":
                a = rng.standard_normal((M, K), dtype=np.float32)
                b = rng.standard_normal((K, Nn), dtype=np.float32)
            else:
                a = rng.uniform(-1, 1, (M, K)).astype(np.float32)
                a[::64] *= 1e3
                b = rng.uniform(-1, 1, (K, Nn)).astype(np.float32)
            A, Bm = vk.Array(gpu, data=a), vk.Array(gpu, data=b)
            rows = np.arange(0, M, sample_every)
            a64 = a[rows].astype(np.float64)
            b64 = b.astype(np.float64)
            c64 = a64 @ b64
            den = np.abs(a64) @ np.abs(b64)
            for mode in ("fp32", "bf16", "bf16x3"):
                tuned = vk.autotune_gemm(gpu, M, Nn, K, precision=mode)
                C = A.matmul(Bm, precision=mode)
                gpu.wait()
                # one untimed product keeps the GPU busy while the timed ones
                # are enqueued behind it, so host latency stays out of the
                # interval between the events
                A.matmul(Bm, precision=mode)
                t = timer.start()
                for _ in range(reps):
                    C = A.matmul(Bm, precision=mode)
                timer.stop(t)
                ms = timer.mean_ms() / reps
                c = np.asarray(C)[rows].astype(np.float64)
                err = np.abs(c - c64)
                rel = err / np.maximum(den, np.finfo(np.float64).tiny)
                out.append({
                    "M": M, "N": Nn, "K": K, "inputs": dist, "mode": mode,
                    "ms": round(ms, 4),
                    "tflops": round(2.0 * M * Nn * K / (ms * 1e-3) / 1e12, 1),
                    "rel_err_max": float(f"{rel.max():.3e}"),
                    "rel_err_mean": float(f"{rel.mean():.3e}"),
                    "normwise_err": float(f"{np.linalg.norm(c - c64) / np.linalg.norm(c64):.3e}"),
                    "config": vk.gemm_configs(mode)[tuned["best"]],
                    "autotune_ms": tuned["ms"],
                })
                log(f"bench: GEMM {out[-1]}")
                del C
            del A, Bm
            gpu.wait()
    return out


def transfer_report(vk, gpu, nbytes: int = 1 << 29, reps: int = 3) -> dict:
    """
    Host <-> device rates (GB/s, PCIe included) for a 512 MiB float32 array:
    DMA from / to its page-locked host mirror, the same copies from / to
    ordinary (pageable) numpy memory for comparison, and the API paths end to
    end: ``Array(gpu, data=x)`` (copy into the mirror + upload, mirror block
    reused from the pinned pool) and ``np.asarray`` of a freshly computed
    array (download).
    """
    from vulkpy_amd import _lib
    n = nbytes // 4
    host = np.random.default_rng(4).standard_normal(n, dtype=np.float32)
    pageable = np.empty(n, dtype=np.float32)
    a = vk.Array(gpu, data=host)
    gpu.wait()
    live = ctypes.c_size_t()
    _lib.call("vkh_host_stats", ctypes.byref(live), None)
    pinned = live.value >= nbytes

    def rate(fn):
        fn()
        gpu.wait()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        gpu.wait()
        return round(nbytes * reps / (time.perf_counter() - t) / 1e9, 2)

    h = gpu._handle
    out = {
        "bytes": nbytes,
        "mirror_pinned": pinned,
        "h2d_pinned": rate(lambda: _lib.call("vkh_upload", h, a._ptr, a._host.ctypes.data, nbytes)),
        "h2d_pageable": rate(lambda: _lib.call("vkh_upload", h, a._ptr, host.ctypes.data, nbytes)),
        "d2h_pinned": rate(lambda: _lib.call("vkh_download", h, a._host.ctypes.data, a._ptr, nbytes)),
        "d2h_pageable": rate(lambda: _lib.call("vkh_download", h, pageable.ctypes.data, a._ptr, nbytes)),
    }

    def create():
        b = vk.Array(gpu, data=host)
        b.wait()

    def readback():
        c = a * 1.0
        np.asarray(c)

    out["array_from_numpy"] = rate(create)
    out["numpy_from_array"] = rate(readback)
    del a
    gpu.wait()
    return out


def cpu_baseline(a, col, row, p, seconds: float = 10.0):
    """Time the oracle's restatement of the reference kernels on a bounded sample."""
    from oracle import oracle as O
    O.load()
    rows = a.shape[0]
    n = a.size
    b = a[::-1].copy()
    suite = [
        ("add_same", 12 * n, lambda: O.binary("add", a, b)),
        ("add_col", 8 * n + 4 * rows, lambda: O.binary_bcast("add", a, col)),
        ("mul_row", 8 * n + 4 * a.shape[1], lambda: O.binary_bcast("mul", a, row)),
        ("sum_axis0", 4 * n + 4 * a.shape[1], lambda: O.reduce_axis("sum", a, 0)),
        ("sum_axis1", 4 * n + 4 * rows, lambda: O.reduce_axis("sum", a, 1)),
        ("max_axis0", 4 * n + 4 * a.shape[1], lambda: O.reduce_axis("maximum", a, 0)),
        ("prod_axis1", 4 * n + 4 * rows, lambda: O.reduce_axis("prod", p, 1)),
    ]
    gen = O.Xoshiro128pp(1 << 14, 1)
    prng_n = n
    suite.append(("xoshiro_float", 4 * prng_n + 32 * (1 << 14), lambda: gen.random(prng_n)))
    for _, _, f in suite:  # warm up
        f()
    total_bytes, t0 = 0, time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < seconds:
        for _, nb, f in suite:
            f()
            total_bytes += nb
        reps += 1
    dt = time.perf_counter() - t0
    # NumPy on the same sample (single-threaded ufuncs; no Xoshiro128++ in NumPy)
    np_suite = [
        (12 * n, lambda: a + b),
        (8 * n + 4 * rows, lambda: a + col),
        (8 * n + 4 * a.shape[1], lambda: a * row),
        (4 * n + 4 * a.shape[1], lambda: a.sum(axis=0)),
        (4 * n + 4 * rows, lambda: a.sum(axis=1)),
        (4 * n + 4 * a.shape[1], lambda: a.max(axis=0)),
        (4 * n + 4 * rows, lambda: p.prod(axis=1)),
    ]
    for _, f in np_suite:
        f()
    np_bytes, np_reps, t1 = 0, 0, time.perf_counter()
    while time.perf_counter() - t1 < seconds / 4:
        for nb, f in np_suite:
            f()
            np_bytes += nb
        np_reps += 1
    np_dt = time.perf_counter() - t1
    return {
        "value": round(total_bytes / dt / 1e9, 3),
        "unit": "GB/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": (f"oracle (C + OpenMP restatement of the reference shaders) on rows 0..{rows - 1} "
                   f"({rows}x{a.shape[1]}) of the same operands: same-shape add, broadcast add/mul, "
                   f"axis sum/max/prod, Xoshiro float; {reps} passes in {dt:.1f} s"),
        "numpy": {"value": round(np_bytes / np_dt / 1e9, 3), "unit": "GB/s", "threads": 1,
                  "sample": (f"NumPy ufuncs on the same sample and ops except Xoshiro; "
                             f"{np_reps} passes in {np_dt:.1f} s")},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--mlp-steps", type=int, default=500,
                    help="training steps per arm of the MLP comparison (0 skips it)")
    ap.add_argument("--mlp-seeds", type=int, default=5,
                    help="weight and data seeds of the MLP seed study (0 skips it)")
    ap.add_argument("--no-gemm", action="store_true", help="skip the GEMM precision report")
    ap.add_argument("--no-c2", action="store_true",
                    help="skip the bandwidth suite (and the CPU baseline that samples it)")
    ap.add_argument("--gemm-pmc", nargs="?", const="bf16", default=None,
                    choices=("fp32", "bf16", "bf16x3"),
                    help="only run five 8192^3 products in this precision (default bf16; for "
                         "counter runs) and exit")
    ap.add_argument("--pmc-kernels", action="store_true",
                    help="only launch the copy and dominant kernels (for counter runs) and exit")
    ap.add_argument("--mlp-trace", type=int, default=0, metavar="S",
                    help="only run S steps of the BF16 and split-BF16 MLP steps with their "
                         "kernels launched one at a time, not as graph replays (for kernel "
                         "traces), and exit")
    ap.add_argument("--mlp-trace-graphs", type=int, default=0, metavar="S",
                    help="only run S steps of the BF16 and split-BF16 MLP steps as graph "
                         "replays (for kernel traces), and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the vulkpy "
              f"path has no multi-GPU mode, see DESIGN.md for how it would shard.",
              file=sys.stderr)
        return 2

    import vulkpy_amd as vk
    gpu = vk.GPU(0)
    timer = Timer(gpu)

    if args.gemm_pmc:
        # for rocprofv3 --pmc runs on the matrix-core kernel alone
        rng = np.random.default_rng(0)
        A = vk.Array(gpu, data=rng.standard_normal((8192, 8192), dtype=np.float32))
        Bm = vk.Array(gpu, data=rng.standard_normal((8192, 8192), dtype=np.float32))
        for _ in range(5):
            A.matmul(Bm, precision=args.gemm_pmc)
        gpu.wait()
        print(json.dumps({"gemm_pmc": f"5 x {args.gemm_pmc} 8192^3"}))
        return 0
    if args.mlp_trace > 0:
        print(json.dumps({"mlp_trace": mlp_compare(vk, gpu, args.mlp_trace, uncaptured=True)}))
        return 0
    if args.mlp_trace_graphs > 0:
        print(json.dumps({"mlp_trace_graphs": mlp_compare(vk, gpu, args.mlp_trace_graphs,
                                                          graphs_only=True)}))
        return 0

    c2 = None
    if not args.no_c2 or args.pmc_kernels:
        c2 = c2_suite(vk, gpu, args, timer)
        if c2 is None:  # --pmc-kernels
            return 0
    mlp = None
    if args.mlp_steps > 0:
        mlp = mlp_compare(vk, gpu, args.mlp_steps, timer=timer)
        if args.mlp_seeds > 0:
            log(f"bench: MLP seed study ({args.mlp_seeds} seeds)")
            mlp["seed_study"] = mlp_seed_study(vk, gpu, range(1, args.mlp_seeds + 1), args.mlp_steps)
    gemm = None
    if not args.no_gemm:
        log("bench: GEMM report")
        gemm = gemm_report(vk, gpu, timer)
    log("bench: host <-> device transfers")
    transfer = transfer_report(vk, gpu)
    log(f"bench: transfer {transfer}")

    cpu = None
    if c2 is not None and not args.no_cpu:
        log("bench: CPU baseline (oracle)")
        rows = 512
        h = c2["host"]
        cpu = cpu_baseline(np.ascontiguousarray(h["a"][:rows]), h["col"][:rows], h["row"],
                           np.ascontiguousarray(h["p"][:rows]), args.cpu_seconds)

    traffic = pmc_traffic()
    out = {
        "metric": METRIC,
        "value": round(c2["value"], 2) if c2 else None,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(c2["ms_per_step"], 3) if c2 else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("C2 bandwidth suite: float32 A (8192,16384) N(0,1) seed 0, broadcast "
                         "operands (8192,1) and (1,16384), same-shape/broadcast/in-place/scalar "
                         "binary ops, sum/prod/max/min/mean over axis 0, 1, None, Xoshiro128++ "
                         "2^20 states -> 2^28 uint32/float/normal"),
            "elements": N,
            "ops_per_step": c2["n_ops"] if c2 else None,
            "bytes_per_step": c2["step_bytes"] if c2 else None,
        },
        "roofline": {
            "kernel": "broadcast binary A (8192,16384) + col (8192,1)",
            "bound": "hbm",
            "achieved": round(c2["achieved"], 1) if c2 else None,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(c2["achieved"] / PEAK_HBM_GBS, 4) if c2 else None,
            "traffic": traffic,
            "algorithmic_bytes": 8 * N + 4 * M_ROWS,
            "avg_launch_ms": round(c2["dom_ms"], 5) if c2 else None,
        },
        "cpu_baseline": cpu,
        "per_op_GBps": c2["per_op"] if c2 else None,
        "per_op_max_ulp": c2["per_op_ulp"] if c2 else None,
        "reduce_tuning": c2["reduce_tuning"] if c2 else None,
        "mlp": mlp,
        "gemm": gemm,
        "transfer": transfer,
    }
    print(json.dumps(out), flush=True)
    return 0


def pmc_traffic():
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over
    ``bench.py --pmc-kernels``, each scaled by the copy kernel's reading), or
    None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "add_col_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes"]
    except (OSError, ValueError, KeyError):
        return None


def c2_suite(vk, gpu, args, timer):
    """The C2 bandwidth suite; returns its measurements (None after the
    --pmc-kernels launches)."""
    log("bench: generating inputs")
    rng = np.random.default_rng(0)
    a_host = rng.standard_normal((M_ROWS, N_COLS), dtype=np.float32)
    col_host = rng.standard_normal((M_ROWS, 1), dtype=np.float32)
    row_host = rng.standard_normal((1, N_COLS), dtype=np.float32)
    b_host = rng.standard_normal((M_ROWS, N_COLS), dtype=np.float32)
    p_host = rng.uniform(0.999, 1.001, (M_ROWS, N_COLS)).astype(np.float32)

    d_host = rng.uniform(-1, 1, (M_ROWS, N_COLS)).astype(np.float32)

    A = vk.Array(gpu, data=a_host)
    B = vk.Array(gpu, data=b_host)
    col = vk.Array(gpu, data=col_host)
    row = vk.Array(gpu, data=row_host)
    P = vk.Array(gpu, data=p_host)
    D = vk.Array(gpu, data=d_host)
    del b_host, d_host
    if args.pmc_kernels:
        # For rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs: a plain copy of
        # known traffic (512 MiB read + 512 MiB written per launch) as the
        # counter scale, then the dominant kernel, 5 launches each.
        from vulkpy_amd import _lib
        dst = vk.Array(gpu, shape=A.shape)
        sh = _lib.shape_array(A.shape)
        gpu.wait()
        for _ in range(5):
            _lib.call("vkh_broadcast_to", gpu._handle, A._ptr, dst._ptr, 2, sh, sh)
        gpu.wait()
        for _ in range(5):
            A + col
        gpu.wait()
        print(json.dumps({"pmc_kernels": ["copy (8192,16384) f32 x5", "add A + col x5"],
                          "copy_bytes": [4 * N, 4 * N], "add_col_bytes": [4 * N + 4 * M_ROWS, 4 * N]}))
        return None
    log("bench: seeding Xoshiro128++ (2^20 states)")
    gen = vk.random.Xoshiro128pp(gpu, size=PRNG_STATES, seed=1)
    ubuf = vk.U32Array(gpu, shape=(PRNG_N,))
    fbuf = vk.Array(gpu, shape=(PRNG_N,))
    gpu.wait()

    # reduction kernel configurations for the suite's three reduction shapes
    # (outside the timed region, like the GEMM tuning)
    log("bench: tuning reductions")
    reduce_tuning = {}
    for name, axis in (("axis0", 0), ("axis1", 1), ("all", None)):
        r = vk.autotune_reduce(gpu, A.shape, axis)[0]
        reduce_tuning[name] = {"family": r["family"], "config": r["config"],
                               "ms_default": r["ms"][0], "ms_best": r["ms"][r["best"]]}
        log(f"bench: reduce {name} {r}")

    # domain operands of the unary functions: Q = |A| + 0.1 > 0 (log, sqrt,
    # pow's base), E = 1 + |A| >= 1 (acosh), D ~ U[-1, 1) (asin, acos, atanh)
    Q = A.abs() + 0.1
    E = A.abs() + 1.0
    ops_in = {"A": A, "B": B, "col": col, "row": row, "P": P, "Q": Q, "D": D, "E": E}
    suite = build_suite(vk, gpu, ops_in, gen, ubuf, fbuf, timer)
    step_bytes = sum(op[1] for op in suite)

    def step():
        for op in suite:
            op[2]()

    log(f"bench: warmup {args.warmup} steps")
    for _ in range(args.warmup):
        step()
    gpu.wait()
    timer.mean_ms()  # drop warmup timings

    n_ops = len(suite)
    log(f"bench: timing {args.steps} steps ({n_ops} ops, {step_bytes / 1e9:.2f} GB per step)")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    gpu.wait()
    dt = time.perf_counter() - t0
    value = step_bytes * args.steps / dt / 1e9

    dom_ms = timer.mean_ms()
    dom_bytes = 8 * N + 4 * M_ROWS
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9

    # per-op GB/s (reported only): HIP events around each op on the library's
    # stream, mean of 3 launches, so host latency is not counted
    per_op = {}
    op_timer = Timer(gpu)
    for name, nb, f, _ in suite:
        gpu.wait()
        for _ in range(3):
            t = op_timer.start()
            f()
            op_timer.stop(t)
        per_op[name] = round(nb / (op_timer.mean_ms() * 1e-3) / 1e9, 1)
    gpu.wait()
    timer.mean_ms()  # drop the dominant op's pairs from this pass

    # per-op error against FP64 on the sample each check returns (reported)
    log("bench: per-op error against FP64")
    per_op_ulp = {}
    for name, _, _, check in suite:
        if callable(check):
            res = check()
            if len(res) == 3 and res[2] is not None:
                got, exact, mag = res
                per_op_ulp[name] = round(float((np.abs(got - exact) / (mag * 2.0 ** -24)).max()), 3)
            else:
                per_op_ulp[name] = round(max_ulp(res[0], res[1]), 3)
    gpu.wait()
    timer.mean_ms()
    del suite, ubuf, fbuf, gen, A, B, P, Q, D, E, ops_in  # free the suite's device buffers
    gpu.wait()
    per_op_ulp.update(prng_check(vk, gpu))
    return {
        "value": value, "ms_per_step": dt / args.steps * 1e3, "n_ops": n_ops,
        "step_bytes": step_bytes, "achieved": achieved, "dom_ms": dom_ms, "per_op": per_op,
        "per_op_ulp": per_op_ulp, "reduce_tuning": reduce_tuning,
        "host": {"a": a_host, "col": col_host, "row": row_host, "p": p_host},
    }


if __name__ == "__main__":
    sys.exit(main())
