"""
Captured training step (:mod:`vulkpy_amd.nn.graph`)
===================================================

``GraphSequence`` runs the whole ``Sequence.train`` step -- forward, loss,
backward, regularizer gradients and optimizer update -- as ONE hipGraph replay
over preallocated buffers:

* forward: one GEMM per Dense with bias and the following ReLU / Sigmoid in
  its epilogue;
* loss: one fused row kernel plus the batch reduction; the loss gradient is
  one kernel;
* backward: per Dense, ``dW = dY^T X``, ``db`` by a column sum, and
  ``dX = dY W`` with the previous activation's derivative in the GEMM
  epilogue;
* update: one kernel per parameter, which first adds the parameter's
  regularizer gradient onto its gradient (writing the sum back, as
  ``Parameter.regular_grad`` leaves it) and then applies the optimizer;
  Adam's beta powers live in device memory and are advanced inside the graph
  (nn/optimizers.py:241-242), once per group of parameters whose powers are
  equal, so no host scalar changes between replays.

In ``precision="fp32"`` every operation is the one ``Sequence.train`` performs
(vulkpy/nn/models.py:55-78 with ``regularize=True``), in the same FP32 order,
so both give identical weights.  ``"bf16"`` runs the GEMMs on BF16 matrix
cores with FP32 accumulation and FP32 master weights.

``train`` returns fresh arrays holding this step's prediction and loss (two
small device copies out of the graph's buffers, which the next replay
overwrites).
"""
from __future__ import annotations

import ctypes
import functools
from typing import Iterable, List, Optional, Tuple

import numpy as np

from vulkpy_amd._lib import ACT, ADAM_MULTI_MAX, GEMM_MODE, AdamParam, call
from vulkpy_amd import vkarray
from vulkpy_amd.vkarray import Array, _DeviceBuffer
from .core import Loss, Module
from .layers import Dense, ReLU, Sigmoid
from .losses import CrossEntropyLoss, HuberLoss, MSELoss, SoftmaxCrossEntropyLoss
from .optimizers import AdaGradState, AdamState, SGDState
from .parameters import Parameter, _has_fused_update
from .regularizers import _fused_penalty

__all__ = ["GraphSequence"]


class _Unit:
    """One Dense layer and the activation that follows it (if any)."""

    def __init__(self, dense: Dense, act: Optional[str]):
        self.dense = dense
        self.act = act


def _units(layers) -> List[_Unit]:
    units: List[_Unit] = []
    for L in layers:
        if isinstance(L, Dense):
            units.append(_Unit(L, None))
        elif isinstance(L, (ReLU, Sigmoid)):
            if not units or units[-1].act is not None:
                raise ValueError("GraphSequence: an activation must follow a Dense layer")
            units[-1].act = "relu" if isinstance(L, ReLU) else "sigmoid"
        else:
            raise ValueError(f"GraphSequence does not support layer {type(L).__name__}")
    if not units:
        raise ValueError("GraphSequence needs at least one Dense layer")
    return units


def _mfma(M: int, N: int, K: int) -> bool:
    """Whether an M x N x K product runs on the matrix-core path (neither
    the skinny nor the short-k vector path)."""
    return min(M, N) > 16 and K > 16


def _packs_a(M: int, N: int, K: int) -> bool:
    """Whether an M x N x K product in BF16 mode packs its operands in a
    layout a BF16 copy can stand in for: the matrix-core path with K a
    multiple of 64."""
    return _mfma(M, N, K) and K % 64 == 0


class _Plan:
    def __init__(self, seq: "GraphSequence", x_shape, y_shape, x: Optional[Array] = None,
                 y: Optional[Array] = None):
        self.key = (tuple(x_shape), tuple(y_shape))
        self.seq = seq
        gpu = seq.gpu
        self.gpu = gpu
        h = gpu._handle
        self.h = h
        # GEMM mode per product family: forward x W^T, input gradient dY W,
        # weight gradient dY^T X
        self.mode_fwd, self.mode_dx, self.mode_dw = (GEMM_MODE[p] for p in _family_precisions(seq.precision))
        units = _units(seq.L)
        self.units = units
        B = int(x_shape[0])
        self.B = B
        if len(x_shape) != 2 or int(x_shape[1]) != units[0].dense.input_dim:
            raise ValueError(f"input shape {x_shape} does not match the first Dense layer")
        for a, b in zip(units, units[1:]):
            if a.dense.output_dim != b.dense.input_dim:
                raise ValueError("consecutive Dense dimensions do not match")
        out_dim = units[-1].dense.output_dim
        if tuple(y_shape) != (B, out_dim):
            raise ValueError(f"target shape {y_shape} != {(B, out_dim)}")
        loss = seq.loss
        if isinstance(loss, SoftmaxCrossEntropyLoss):
            self.loss_kind = "sce"
        elif isinstance(loss, CrossEntropyLoss):
            self.loss_kind = "ce"
        elif isinstance(loss, MSELoss):
            self.loss_kind = "mse"
        elif isinstance(loss, HuberLoss):
            self.loss_kind = "huber"
        else:
            raise ValueError(f"GraphSequence does not support loss {type(loss).__name__}")
        self.mean = loss.scale_backward is not None
        self.scale = float(np.float32(1 / B)) if self.mean else 1.0

        # static buffers
        self.x = Array(gpu, shape=(B, units[0].dense.input_dim))
        self.y = Array(gpu, shape=(B, out_dim))
        self.hs = [Array(gpu, shape=(B, u.dense.output_dim)) for u in units]
        self.ds = [Array(gpu, shape=(B, u.dense.output_dim)) for u in units]
        self.sm = Array(gpu, shape=(B, out_dim)) if self.loss_kind == "sce" else None
        self.g_out = Array(gpu, shape=(B, out_dim)) if units[-1].act else None
        self.L_rows = Array(gpu, shape=(B,))
        self.loss = Array(gpu, shape=())
        # BF16 mode: BF16 copies of the activations hs[i] and input gradients
        # ds[i] that the next product would pack, written by the epilogue of
        # the product that makes them (c_bf16) and taken by the consumer as
        # its packed operand (vkh_gemm_bf16a), which skips that packing pass
        # with the same bits
        self.hs_bf: List[Optional[_DeviceBuffer]] = [None] * len(units)
        self.ds_bf: List[Optional[_DeviceBuffer]] = [None] * len(units)
        # ... and their transposes hs[i]^T and ds[i]^T, written by the same
        # products (c_bf16t, one 64-byte run of C^T per column of each 32-row
        # tile) for the weight-gradient products dW = ds[i]^T x, which take
        # them as packed op(A) and op(B)^T (vkh_gemm_bf16io) instead of
        # packing with a transposing pass; only where the producer runs on
        # the matrix cores, whose epilogue writes C^T coalesced
        self.hs_bft: List[Optional[_DeviceBuffer]] = [None] * len(units)
        self.ds_bft: List[Optional[_DeviceBuffer]] = [None] * len(units)
        # a last layer with one output after a hidden layer: its backward and
        # the hidden layer's bias gradient in one pass (vkh_dense1_backward)
        last = units[-1].dense
        self.fuse_last = False
        if len(units) > 1 and last.output_dim == 1:
            ok = ctypes.c_int(0)
            call("vkh_dense1_backward_supported", gpu._handle, B, last.input_dim, ctypes.byref(ok))
            self.fuse_last = bool(ok.value)
        bf16 = GEMM_MODE["bf16"]
        rows = -(-B // 256) * 256
        if self.mode_fwd == bf16:
            for i in range(len(units) - 1):   # consumer: forward product of unit i + 1
                k = units[i].dense.output_dim
                if _packs_a(B, units[i + 1].dense.output_dim, k):
                    self.hs_bf[i] = _DeviceBuffer(gpu, rows * k, np.uint16)
        if self.mode_dx == bf16:
            for i in range(1, len(units) - 1):  # consumer: dX product of unit i
                k = units[i].dense.output_dim
                if _packs_a(B, units[i].dense.input_dim, k):
                    self.ds_bf[i] = _DeviceBuffer(gpu, rows * k, np.uint16)
        if self.mode_dw == bf16:
            for i in range(len(units) - 1):
                d, nxt = units[i].dense, units[i + 1].dense
                trows = -(-d.output_dim // 256) * 256
                # hs[i]^T: forward product of unit i -> op(B)^T of dW of unit i + 1
                if (nxt.w.grad is not None and _mfma(B, d.output_dim, d.input_dim)
                        and _packs_a(nxt.output_dim, d.output_dim, B)):
                    self.hs_bft[i] = _DeviceBuffer(gpu, trows * B, np.uint16)
                # ds[i]^T: dX product of unit i + 1 (or the fused last-layer
                # backward) -> op(A) of dW of unit i
                producer = _mfma(B, d.output_dim, nxt.output_dim) or (
                    self.fuse_last and i == len(units) - 2)
                if d.w.grad is not None and producer and _packs_a(d.output_dim, d.input_dim, B):
                    self.ds_bft[i] = _DeviceBuffer(gpu, trows * B, np.uint16)
        # parameters, their optimizer states, regularizer terms and device
        # beta powers
        self.params: List[Parameter] = []
        self.reg = {}
        for u in units:
            for p in (u.dense.w, u.dense.b):
                if p.grad is None:
                    continue
                st = p.opt_state
                if not _has_fused_update(st) or not isinstance(st, (AdamState, SGDState, AdaGradState)):
                    raise ValueError(f"GraphSequence does not support optimizer state {type(st).__name__}")
                self.params.append(p)
                terms = (0, 0.0, 0.0)
                if seq.regularize and p.R is not None:
                    if not _fused_penalty(p.R):
                        raise ValueError(f"GraphSequence does not support regularizer {type(p.R).__name__}")
                    terms = p.R.terms()
                self.reg[id(p)] = terms
        # Adam states whose beta powers evolve together (same betas, same
        # powers so far) share one FP64 device pair, advanced once per step
        self.bpow = {}
        self.bpow_groups: List[Tuple[_DeviceDoubles, float, float]] = []
        groups = {}
        for p in self.params:
            st = p.opt_state
            if isinstance(st, AdamState):
                key = (float(st.opt.beta1), float(st.opt.beta2), float(st.beta1t), float(st.beta2t))
                if key not in groups:
                    buf = _DeviceDoubles(gpu, np.array([st.beta1t, st.beta2t], dtype=np.float64))
                    groups[key] = buf
                    self.bpow_groups.append((buf, key[0], key[1]))
                self.bpow[id(st)] = groups[key]
        # the inputs copied into the step's buffers last time, and their jobs
        self._src = {"x": (None, None), "y": (None, None)}
        self._events = None  # (product key, start, end) timer events of a timed tuning pass
        gpu.wait()
        self.graph = None
        self._capture(x, y)

    # ---- the step -------------------------------------------------------------------
    def _enqueue(self, tune: bool = False, update: bool = True):
        """Issue the step's kernels.  tune: a pass that times every product's
        kernel configurations with its own operands and epilogue
        (vkh_gemm_autotune_ext) and records the products' shapes; update
        False (implied by tune): stop before the optimizer update."""
        h, B = self.h, self.B
        units = self.units

        def gemm(*args):  # the arguments of vkh_gemm_bf16io_ext
            key = tuple(args[1:7])  # mode, trans_a, trans_b, M, N, K
            if tune:
                call("vkh_gemm_autotune_ext", *args, 3, None, None)
                if _mfma(*key[3:]) and key not in self._products:
                    self._products.append(key)
                    self._takes_colsum[key] = args[26] is not None
            elif self._events is not None:  # timed pass of the step tuning
                e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
                call("vkh_timer_record", h, ctypes.byref(e0))
                call("vkh_gemm_bf16io_ext", *args)
                call("vkh_timer_record", h, ctypes.byref(e1))
                self._events.append((key, e0, e1))
            else:
                call("vkh_gemm_bf16io_ext", *args)

        # forward
        src, src_bf = self.x, None
        for i, (u, hout) in enumerate(zip(units, self.hs)):
            d = u.dense
            hb, hbt = self.hs_bf[i], self.hs_bft[i]
            store = self.h_store[i]
            gemm(h, self.mode_fwd, 0, 1, B, d.output_dim, d.input_dim, 1.0, src._ptr,
                 d.input_dim, d.w.value._ptr, d.input_dim, 0.0, hout._ptr if store else None, d.output_dim,
                 d.b.value._ptr, ACT[u.act], hb.ptr if hb else None, None, None, 0, 0,
                 src_bf.ptr if src_bf else None, None, hbt.ptr if hbt else None, B, None, int(store))
            src, src_bf = hout, hb
        out = self.hs[-1]
        O = units[-1].dense.output_dim
        # loss rows, batch reduction and gradient w.r.t. the network output
        gl = self.g_out if self.g_out is not None else self.ds[-1]
        if self.loss_kind == "sce":
            call("vkh_softmax_cross_entropy", h, out._ptr, self.y._ptr, self.sm._ptr, self.L_rows._ptr, B, O)
            call("vkh_loss_grad", h, 0, self.sm._ptr, self.y._ptr, gl._ptr, B * O, self.scale)
        elif self.loss_kind == "ce":
            call("vkh_cross_entropy", h, out._ptr, self.y._ptr, self.L_rows._ptr, B, O)
            call("vkh_cross_entropy_backward", h, out._ptr, self.y._ptr, gl._ptr, B * O, self.scale)
        else:
            # loss rows and gradient in one kernel (vkh_mse_rows / vkh_huber_rows
            # and vkh_loss_grad, the same bits)
            kind = 1 if self.loss_kind == "mse" else 2
            call("vkh_loss_rows_grad", h, kind, out._ptr, self.y._ptr, self.L_rows._ptr, gl._ptr, B, O,
                 self.scale)
        mul = float(np.float32(1 / B)) if self.mean else 1.0
        call("vkh_reduce_axis", h, 0, self.L_rows._ptr, self.loss._ptr, 1, B, 1, 0, mul, 1.0)
        if self.g_out is not None:
            call("vkh_activation_backward", h, ACT[units[-1].act], out._ptr, self.g_out._ptr,
                 self.ds[-1]._ptr, B * O)
        # backward
        last = len(units) - 1
        for i in range(last, -1, -1):
            d = units[i].dense
            dy = self.ds[i]
            xin = self.x if i == 0 else self.hs[i - 1]
            fused = self.fuse_last and i == last
            for p, is_w in ((d.w, True), (d.b, False)):
                if p.grad is None:
                    continue
                if (is_w and fused) or (not is_w and self.fuse_last and i == last - 1):
                    continue  # taken by vkh_dense1_backward
                if not is_w and (i + 1) in self.colsum:
                    continue  # taken by the next unit's input-gradient product
                # the regularizer term is added by the optimizer kernel
                beta = 0.0
                if is_w:
                    dyt, xt = self.ds_bft[i], (self.hs_bft[i - 1] if i > 0 else None)
                    gemm(h, self.mode_dw, 1, 0, d.output_dim, d.input_dim, B, 1.0, dy._ptr,
                         d.output_dim, xin._ptr, d.input_dim, beta, p.grad._ptr, d.input_dim,
                         None, 0, None, None, None, 0, 0, dyt.ptr if dyt else None, xt.ptr if xt else None,
                         None, 0, None, 1)
                else:
                    call("vkh_bias_grad", h, dy._ptr, p.grad._ptr, B, d.output_dim, beta)
            if fused:
                prev = units[i - 1]
                out_bf, out_bft = self.ds_bf[i - 1], self.ds_bft[i - 1]
                pb = prev.dense.b.grad
                call("vkh_dense1_backward", h, dy._ptr, d.w.value._ptr, self.hs[i - 1]._ptr,
                     ACT[prev.act], B, d.input_dim, self.ds[i - 1]._ptr if self.d1_store else None,
                     out_bf.ptr if out_bf else None, out_bft.ptr if out_bft else None, B,
                     d.w.grad._ptr if d.w.grad is not None else None,
                     pb._ptr if pb is not None else None)
            elif i > 0:
                prev = units[i - 1]
                out_bf, dy_bf, out_bft = self.ds_bf[i - 1], self.ds_bf[i], self.ds_bft[i - 1]
                # ReLU' from the BF16 copy of h when there is one (the same
                # factor: the copy keeps the sign and zero-ness of h)
                aux16 = self.hs_bf[i - 1] if prev.act == "relu" else None
                aux32 = self.hs[i - 1]._ptr if prev.act and aux16 is None else None
                # with the bias gradient of unit i - 1 from its epilogue
                db = prev.dense.b.grad._ptr if i in self.colsum else None
                store = int(self.colsum[i]) if i in self.colsum else 1
                gemm(h, self.mode_dx, 0, 0, B, d.input_dim, d.output_dim, 1.0, dy._ptr,
                     d.output_dim, d.w.value._ptr, d.input_dim, 0.0, self.ds[i - 1]._ptr, d.input_dim, None, 0,
                     out_bf.ptr if out_bf else None, aux32, aux16.ptr if aux16 else None, d.input_dim,
                     ACT[prev.act], dy_bf.ptr if dy_bf else None, None, out_bft.ptr if out_bft else None, B,
                     db, store)
        if tune or not update:
            return
        # optimizer: beta powers once per group, then one kernel per
        # parameter (regularizer term, gradient write-back and update)
        for buf, beta1, beta2 in self.bpow_groups:
            call("vkh_adam_pow_advance", h, buf.ptr, beta1, beta2)
        # Adam parameters: up to ADAM_MULTI_MAX a launch
        adam = []
        for p in self.params:
            st = p.opt_state
            if isinstance(st, AdamState):
                o = st.opt
                kind, l1, s2 = self.reg[id(p)]
                adam.append(AdamParam(p.value._ptr, p.grad._ptr, st.m._ptr, st.v._ptr, p.value.size,
                                      float(o.lr), float(o.eps), float(o.beta1), float(o.beta2),
                                      self.bpow[id(st)].ptr, kind, l1, s2))
        for k in range(0, len(adam), ADAM_MULTI_MAX):
            part = adam[k:k + ADAM_MULTI_MAX]
            call("vkh_adam_device_multi", h, (AdamParam * len(part))(*part), len(part))
        for p in self.params:
            st = p.opt_state
            n = p.value.size
            kind, l1, s2 = self.reg[id(p)]
            if isinstance(st, AdamState):
                continue
            elif isinstance(st, SGDState):
                call("vkh_sgd", h, p.value._ptr, p.grad._ptr, n, float(st.opt.lr), kind, l1, s2)
            else:
                call("vkh_adagrad", h, p.value._ptr, p.grad._ptr, st.h._ptr, n,
                     float(st.opt.lr), float(st.opt.eps), kind, l1, s2)

    def _gemm_shapes(self):
        """(mode, trans_a, trans_b, M, N, K) of every product _enqueue issues"""
        B = self.B
        for i, u in enumerate(self.units):
            d = u.dense
            yield self.mode_fwd, 0, 1, B, d.output_dim, d.input_dim      # forward x W^T
            yield self.mode_dw, 1, 0, d.output_dim, d.input_dim, B       # dW = dY^T X
            if i > 0:
                yield self.mode_dx, 0, 0, B, d.input_dim, d.output_dim   # dX = dY W

    def _reduce_shapes(self):
        """(prev, axis, post) of every axis reduction _enqueue issues"""
        B = self.B
        yield 1, B, 1                                     # batch loss
        for d in sorted({u.dense.output_dim for u in self.units if u.dense.b.grad is not None}):
            yield 1, B, d                                 # bias gradient (vkh_bias_grad)

    def _plan_colsums(self, check: bool = True):
        """Input-gradient products that also give the previous layer's bias
        gradient (vkh_gemm_bf16io_colsum), decided once the GEMM
        configurations are chosen; and whether their FP32 output is read at
        all (not when every consumer takes its BF16 copies)."""
        self.colsum = {}
        B, units = self.B, self.units
        last = len(units) - 1
        self.h_store = [self._h_fp32_read(i) for i in range(len(units))]
        # the fused last-layer backward also gives the hidden layer's bias
        # gradient, so its FP32 output is only for the products
        self.d1_store = self.fuse_last and self._dy_fp32_read(last - 1)
        for i in range(1, len(units)):
            if self.fuse_last and i == last:
                continue
            d, prev = units[i].dense, units[i - 1].dense
            if prev.b.grad is None:
                continue
            if check:  # (without: as if every product took the column sums, for the tuning pass)
                ok = ctypes.c_int(0)
                call("vkh_gemm_colsum_supported", self.h, self.mode_dx, 0, 0, B, d.input_dim, d.output_dim,
                     ctypes.byref(ok))
                if not ok.value:
                    continue
            self.colsum[i] = self._dy_fp32_read(i - 1)

    def _h_fp32_read(self, i: int) -> bool:
        """Whether anything reads the FP32 output hs[i] of unit i's forward
        product: the prediction and the fused last-layer backward do; the next
        unit's forward, weight-gradient and input-gradient products do unless
        they take BF16 copies (ReLU' from the row copy, no other activation)."""
        last = len(self.units) - 1
        if i == last or (self.fuse_last and i == last - 1):
            return True
        nxt = self.units[i + 1].dense
        fwd_bf = self.hs_bf[i] is not None
        dw_bf = nxt.w.grad is None or self.hs_bft[i] is not None
        act = self.units[i].act
        aux_bf = act is None or (act == "relu" and self.hs_bf[i] is not None)
        return not (fwd_bf and dw_bf and aux_bf)

    def _step_tune(self, rounds: int = 2, passes: int = 2):
        """Second round of the GEMM tuning, in the step's own context.  A
        product alone and repeated keeps its operands in cache; in the step
        its neighbours' traffic decides too.  So the forward and backward
        pass (no update) runs with every product in each of its candidate
        configurations in turn, timing each product's call between events,
        `rounds` times over; each product keeps the configuration with the
        lowest median.  A product whose epilogue takes the column sums
        only tries the configurations that can."""
        h = self.h
        cands = {}
        for key in self._products:
            n = ctypes.c_int(0)
            call("vkh_gemm_configs", key[0], ctypes.byref(n))
            allowed = []
            for c in range(n.value):
                call("vkh_gemm_set_config", h, *key, c)
                if self._takes_colsum[key]:
                    ok = ctypes.c_int(0)
                    call("vkh_gemm_colsum_supported", h, *key, ctypes.byref(ok))
                    if not ok.value:
                        continue
                allowed.append(c)
            cands[key] = allowed or [0]
        times = {key: {c: [] for c in cands[key]} for key in cands}
        trials = max((len(v) for v in cands.values()), default=0)
        ms = ctypes.c_float()
        for _ in range(rounds):
            for t in range(trials):
                chosen = {key: cands[key][t % len(cands[key])] for key in cands}
                for key, c in chosen.items():
                    call("vkh_gemm_set_config", h, *key, c)
                self._plan_colsums()
                self._enqueue(update=False)
                self._events = []
                try:
                    for _ in range(passes):
                        self._enqueue(update=False)
                    self.gpu.wait()
                    for key, e0, e1 in self._events:
                        call("vkh_event_elapsed_ms", e0, e1, ctypes.byref(ms))
                        if key in chosen:
                            times[key][chosen[key]].append(ms.value)
                finally:
                    for _, e0, e1 in self._events:
                        call("vkh_event_destroy", e0)
                        call("vkh_event_destroy", e1)
                    self._events = None
        for key, per in times.items():
            best = min(per, key=lambda c: float(np.median(per[c])) if per[c] else float("inf"))
            call("vkh_gemm_set_config", h, *key, best)

    def _dy_fp32_read(self, p: int) -> bool:
        """Whether the products of unit p read its FP32 output gradient ds[p]:
        its weight gradient does unless it takes the transposed BF16 copy, its
        input gradient (p > 0) unless it takes the BF16 row copy."""
        bf16 = GEMM_MODE["bf16"]
        d = self.units[p].dense
        dw_bf = (d.w.grad is None or (self.mode_dw == bf16 and self.ds_bft[p] is not None
                                      and self.B % 64 == 0))
        dx_bf = p == 0 or (self.mode_dx == bf16 and self.ds_bf[p] is not None and d.output_dim % 64 == 0)
        return not (dw_bf and dx_bf)

    def _capture(self, x: Optional[Array] = None, y: Optional[Array] = None):
        # the graph's GEMM packing region for every product of the step
        for mode, ta, tb, M, N, K in self._gemm_shapes():
            call("vkh_gemm_reserve", self.h, mode, M, N, K)
        if self.seq.autotune:
            for prev, K, post in self._reduce_shapes():
                # only configurations with the default summation order, so
                # the step's results never depend on the timings
                call("vkh_reduce_autotune", self.h, 1, prev, K, post, 3, 1, None, None)
            # GEMM configurations: a pass of the step (on the first batch when
            # given, no optimizer update) that times each product with its own
            # operands and epilogue, whose cost differs between tile shapes
            for name, src, dst in (("x", x, self.x), ("y", y, self.y)):
                if src is not None:
                    call("vkh_copy", self.h, dst._ptr, src._ptr, dst.buffer.nbytes)
                    self._src[name] = (src, src.job)
            self._plan_colsums(check=False)
            self._products, self._takes_colsum = [], {}
            self._enqueue(tune=True)
            self._step_tune()
        self._plan_colsums()
        self.gpu.wait()
        if not self.seq.capture:
            return
        call("vkh_graph_begin", self.h)
        vkarray._capture_begin()
        try:
            try:
                self._enqueue()
            except BaseException:
                ex = ctypes.c_void_p()
                from vulkpy_amd import _lib
                _lib.lib.vkh_graph_end(self.h, ctypes.byref(ex))
                if ex.value:
                    _lib.lib.vkh_graph_destroy(ex.value)
                raise
            ex = ctypes.c_void_p()
            call("vkh_graph_end", self.h, ctypes.byref(ex))
            self.graph = ex.value
        finally:
            vkarray._capture_end()

    def run(self, x: Array, y: Array) -> Tuple[Array, Array]:
        h = self.h
        # an input that is the array copied last time, unchanged since (its
        # job is the same), is in the step's buffer already
        for name, src, dst in (("x", x, self.x), ("y", y, self.y)):
            last, job = self._src[name]
            if src is not last or src.job is not job:
                call("vkh_copy", h, dst._ptr, src._ptr, dst.buffer.nbytes)
                self._src[name] = (src, src.job)
        if self.graph is None:
            self._enqueue()
        else:
            call("vkh_graph_launch", h, self.graph)
        for p in self.params:
            st = p.opt_state
            if isinstance(st, AdamState):
                st.beta1t *= st.opt.beta1
                st.beta2t *= st.opt.beta2
            p.value._set_job()
            p.grad._set_job()
            for name in ("m", "v", "h"):
                arr = getattr(st, name, None)
                if isinstance(arr, Array):
                    arr._set_job()
        # every array the step writes gets the new ticket (and a stale host
        # mirror is dropped)
        for a in self.hs + self.ds + [self.loss, self.L_rows, self.sm, self.g_out]:
            if a is not None:
                a._set_job()
        # fresh arrays, like the reference's train (nn/models.py:55-78): the
        # next replay overwrites the graph's own output buffers
        pred, loss = Array(self.gpu, shape=self.hs[-1].shape), Array(self.gpu, shape=())
        call("vkh_copy", h, pred._ptr, self.hs[-1]._ptr, pred.buffer.nbytes)
        call("vkh_copy", h, loss._ptr, self.loss._ptr, loss.buffer.nbytes)
        pred._set_job([self.hs[-1]])
        loss._set_job([self.loss])
        return pred, loss

    def __del__(self):
        try:
            if self.graph:
                from vulkpy_amd import _lib
                graph, self.graph = self.graph, None
                # not while another step is being captured
                if not vkarray._after_capture(functools.partial(_lib.lib.vkh_graph_destroy, graph)):
                    _lib.lib.vkh_graph_destroy(graph)
        except Exception:
            pass


_FAMILIES = ("forward", "input_grad", "weight_grad")


def _family_precisions(precision) -> Tuple[str, str, str]:
    """GEMM precision of the forward, input-gradient and weight-gradient
    products from ``precision`` (one name for all three, or a dict by family)"""
    if isinstance(precision, dict):
        unknown = set(precision) - set(_FAMILIES)
        if unknown:
            raise ValueError(f"unknown product families {sorted(unknown)}")
        out = tuple(precision.get(f, "bf16") for f in _FAMILIES)
    else:
        out = (precision,) * 3
    for p in out:
        if p not in GEMM_MODE:
            raise ValueError(f"unknown precision {p!r}")
    return out


class _DeviceDoubles:
    """Small FP64 device buffer (Adam beta powers)."""

    def __init__(self, gpu, values: np.ndarray):
        from vulkpy_amd.vkarray import _DeviceBuffer
        self._host = np.ascontiguousarray(values, dtype=np.float64)
        self._buf = _DeviceBuffer(gpu, self._host.size, np.float64)
        call("vkh_upload", gpu._handle, self._buf.ptr, self._host.ctypes.data, self._host.nbytes)
        gpu.wait()

    @property
    def ptr(self):
        return self._buf.ptr


class GraphSequence:
    """
    Sequential model whose training step is one captured hipGraph

    Parameters
    ----------
    layers : iterable of vulkpy_amd.nn.Module
        Dense layers, each optionally followed by ReLU or Sigmoid.
    loss : vulkpy_amd.nn.Loss
        SoftmaxCrossEntropyLoss, CrossEntropyLoss, MSELoss or HuberLoss.
    precision : {"bf16", "fp32", "bf16x3"} or dict
        Product precision of the GEMMs (FP32 accumulation and FP32 master
        weights in every mode); a dict sets it per product family, keys
        ``"forward"``, ``"input_grad"`` and ``"weight_grad"`` (missing keys:
        ``"bf16"``).
    regularize : bool
        Add each parameter's regularizer gradient before the update
        (``Parameter.regular_grad``).  The reference's ``Sequence.train``
        never does (vulkpy/nn/models.py:55-78).
    autotune : bool
        Before capturing, time every GEMM kernel configuration on each
        product of the step with its own operands and epilogue (a pass of the
        step on the first batch without the optimizer update, through
        ``vkh_gemm_autotune_ext``) and the reduction configurations of the
        bias gradients and the batch loss that keep the default summation
        order (``vulkpy_amd.autotune_reduce(..., same_order=True)``), and use
        the fastest.  Neither choice changes a result.  The choice is kept on
        the GPU context, so eager calls of the same shapes use it too.
    capture : bool
        Capture the step as a hipGraph (default).  With False every ``train``
        call issues the same kernels, with the same arguments and results,
        one launch at a time: for profilers and debuggers that do not follow
        graph replays.

    Notes
    -----
    The graph is captured on the first ``train`` call for a given pair of
    input / target shapes and replayed afterwards.
    """

    def __init__(self, layers: Iterable[Module], loss: Loss, *,
                 precision: str = "bf16", regularize: bool = True, autotune: bool = False,
                 capture: bool = True):
        _family_precisions(precision)  # validates
        self.L: Tuple[Module, ...] = tuple(layers)
        self.loss = loss
        self.precision = precision
        self.regularize = regularize
        self.autotune = autotune
        self.capture = capture
        self._plan: Optional[_Plan] = None
        units = _units(self.L)
        self.gpu = units[0].dense.w.value._gpu

    def train(self, x: Array, y: Array) -> Tuple[Array, Array]:
        """
        One training step; returns (prediction, loss) as fresh arrays.
        """
        key = (tuple(x.shape), tuple(y.shape))
        if self._plan is None or self._plan.key != key:
            self._plan = None
            self._plan = _Plan(self, x.shape, y.shape, x, y)
        return self._plan.run(x, y)

    def predict(self, x: Array, y: Optional[Array] = None):
        """Forward (and loss) through the layers' ordinary path."""
        for L in self.L:
            x = L(x)
        if y is None:
            return x
        return x, self.loss(x, y)
