#!/usr/bin/env python3
"""
Bandwidth benchmark of the vulkpy array path on one MI355X.

Workload (one "step"): the 8192 x 16384 float32 suite -- A ~ N(0, 1) from
numpy.random.default_rng(0), broadcast operands col (8192, 1) and row
(1, 16384), a second full operand B, and P ~ U[0.999, 1.001] for products:

  * same-shape binary ops A (+ - * /) B                       12 N bytes each
  * broadcast ops A (+ - * / max min) col and ... row         8 N + |operand|
  * in-place broadcast ops A (+= -= *= /=) col / row          8 N + |operand|
  * scalar op A * 2                                           8 N
  * sum / maximum / minimum / mean over axis 0, 1 and None,
    prod of P over axis 0, 1 and None                         4 N + |output|
  * Xoshiro128++ with 2^20 states: 2^28 uint32, float, normal 4 * 2^28 + 32 * 2^20

value = algorithmic bytes of all ops / wall time of the timed steps (GB/s,
inputs resident in HBM).  The dominant kernel (broadcast A + col) is also
timed per launch with HIP events on the library's stream for the roofline.
The CPU baseline times the oracle (C + OpenMP, the reference's algorithms) on
the first 512 rows of the same operands.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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


def build_suite(vk, gpu, A, B, col, row, P, rng_gen, ubuf, fbuf, timer):
    """List of (name, algorithmic bytes, callable) for one step."""
    ops = []
    colb, rowb = 4 * M_ROWS, 4 * N_COLS

    def add(name, nbytes, fn):
        ops.append((name, nbytes, fn))

    for o, f in (("add", lambda x, y: x + y), ("sub", lambda x, y: x - y),
                 ("mul", lambda x, y: x * y), ("div", lambda x, y: x / y)):
        add(f"{o}_same", 12 * N, (lambda f: lambda: f(A, B))(f))
    for o, f in (("add", lambda x, y: x + y), ("sub", lambda x, y: x - y),
                 ("mul", lambda x, y: x * y), ("div", lambda x, y: x / y),
                 ("max", lambda x, y: x.max(y)), ("min", lambda x, y: x.min(y))):
        if o == "add":
            def timed(f=f):
                t = timer.start()
                r = f(A, col)
                timer.stop(t)
                return r
            add("add_col", 8 * N + colb, timed)
        else:
            add(f"{o}_col", 8 * N + colb, (lambda f: lambda: f(A, col))(f))
        add(f"{o}_row", 8 * N + rowb, (lambda f: lambda: f(A, row))(f))

    def inplace():
        A.__iadd__(col)
        A.__isub__(col)
        A.__imul__(row)
        A.__itruediv__(row)
    add("inplace_bcast_x4", 4 * (8 * N) + 2 * colb + 2 * rowb, inplace)
    add("mul_scalar", 8 * N, lambda: A * 2.0)
    for f in ("sum", "maximum", "minimum", "mean"):
        add(f"{f}_axis0", 4 * N + 4 * N_COLS, (lambda f: lambda: getattr(A, f)(axis=0))(f))
        add(f"{f}_axis1", 4 * N + 4 * M_ROWS, (lambda f: lambda: getattr(A, f)(axis=1))(f))
        add(f"{f}_all", 4 * N + 4, (lambda f: lambda: getattr(A, f)())(f))
    add("prod_axis0", 4 * N + 4 * N_COLS, lambda: P.prod(axis=0))
    add("prod_axis1", 4 * N + 4 * M_ROWS, lambda: P.prod(axis=1))
    add("prod_all", 4 * N + 4, lambda: P.prod())
    prng_bytes = 4 * PRNG_N + 32 * PRNG_STATES
    add("xoshiro_uint32", prng_bytes, lambda: rng_gen.randint(buffer=ubuf))
    add("xoshiro_float", prng_bytes, lambda: rng_gen.random(buffer=fbuf))
    add("xoshiro_normal", prng_bytes, lambda: rng_gen.normal(buffer=fbuf))
    return ops


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
    return {
        "value": round(total_bytes / dt / 1e9, 3),
        "unit": "GB/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": (f"oracle (C + OpenMP restatement of the reference shaders) on rows 0..{rows - 1} "
                   f"({rows}x{a.shape[1]}) of the same operands: same-shape add, broadcast add/mul, "
                   f"axis sum/max/prod, Xoshiro float; {reps} passes in {dt:.1f} s"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the vulkpy "
              f"path has no multi-GPU mode, see DESIGN.md for how it would shard.",
              file=sys.stderr)
        return 2

    import vulkpy_amd as vk
    gpu = vk.GPU(0)

    log("bench: generating inputs")
    rng = np.random.default_rng(0)
    a_host = rng.standard_normal((M_ROWS, N_COLS), dtype=np.float32)
    col_host = rng.standard_normal((M_ROWS, 1), dtype=np.float32)
    row_host = rng.standard_normal((1, N_COLS), dtype=np.float32)
    b_host = rng.standard_normal((M_ROWS, N_COLS), dtype=np.float32)
    p_host = rng.uniform(0.999, 1.001, (M_ROWS, N_COLS)).astype(np.float32)

    A = vk.Array(gpu, data=a_host)
    B = vk.Array(gpu, data=b_host)
    col = vk.Array(gpu, data=col_host)
    row = vk.Array(gpu, data=row_host)
    P = vk.Array(gpu, data=p_host)
    del b_host
    log("bench: seeding Xoshiro128++ (2^20 states)")
    gen = vk.random.Xoshiro128pp(gpu, size=PRNG_STATES, seed=1)
    ubuf = vk.U32Array(gpu, shape=(PRNG_N,))
    fbuf = vk.Array(gpu, shape=(PRNG_N,))
    gpu.wait()

    timer = Timer(gpu)
    suite = build_suite(vk, gpu, A, B, col, row, P, gen, ubuf, fbuf, timer)
    step_bytes = sum(nb for _, nb, _ in suite)

    def step():
        for _, _, f in suite:
            f()

    log(f"bench: warmup {args.warmup} steps")
    for _ in range(args.warmup):
        step()
    gpu.wait()
    timer.mean_ms()  # drop warmup timings

    log(f"bench: timing {args.steps} steps ({len(suite)} ops, {step_bytes / 1e9:.2f} GB per step)")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    gpu.wait()
    dt = time.perf_counter() - t0
    value = step_bytes * args.steps / dt / 1e9

    dom_ms = timer.mean_ms()
    dom_bytes = 8 * N + 4 * M_ROWS
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9

    # per-op GB/s from one extra, individually synchronised pass (reported only)
    per_op = {}
    for name, nb, f in suite:
        gpu.wait()
        t = time.perf_counter()
        f()
        gpu.wait()
        per_op[name] = round(nb / (time.perf_counter() - t) / 1e9, 1)

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle)")
        rows = 512
        cpu = cpu_baseline(np.ascontiguousarray(a_host[:rows]), col_host[:rows], row_host,
                           np.ascontiguousarray(p_host[:rows]), args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
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
            "ops_per_step": len(suite),
            "bytes_per_step": step_bytes,
        },
        "roofline": {
            "kernel": "broadcast binary A (8192,16384) + col (8192,1)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": None,
            "avg_launch_ms": round(dom_ms, 5),
        },
        "cpu_baseline": cpu,
        "per_op_GBps": per_op,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
