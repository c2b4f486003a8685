"""
torch.autograd wrappers over the HIP kernels (:mod:`vulkpy_amd.torch_autograd`)
===============================================================================

Differentiable functions on float32 torch tensors that live on a ROCm GPU.
Forward and backward are the library's hand-written kernels, launched on
torch's current stream through a context bound to that stream
(``vkh_gpu_create_on_stream``); torch provides the tensors and the autograd
tape only.

=================================  =====================================  ==============================
function                           forward                                backward
=================================  =====================================  ==============================
``dense(x, w, b, precision)``      ``x @ w.T + b`` (nn/layers.py:94-101)   ``dx = dy @ w``, ``dw = dy.T @ x``,
                                                                           ``db = dy.sum(0)``
                                                                           (nn/layers.py:126-141)
``relu(x)``, ``sigmoid(x)``        nn/layers.py:185, 240-244              nn/layers.py:207-210, 264-267
``softmax(x)``                     row softmax (nn/layers.py:297-300)     exact Jacobian-vector product
``softmax_cross_entropy(x, t)``    nn/losses.py:200-227 + reduction       ``softmax(x) - t`` (nn/losses.py:249)
``mse(x, t)``                      nn/losses.py:294-296 + reduction       ``2 (x - t)`` (nn/losses.py:318-320)
``huber(x, t)``                    nn/losses.py:365-369 + reduction       exact derivative of the forward
=================================  =====================================  ==============================

Losses reduce per-sample row sums by ``"mean"`` (default) or ``"sum"`` over the
batch like ``ReduceLoss`` (nn/losses.py:42-86), and their gradients are scaled
by 1/B for the mean.  Two backward formulas differ from the reference's layer
code on purpose, so that the gradients are those of the forward functions:
Softmax (the reference multiplies by the Jacobian diagonal ``y (1 - y)``,
nn/layers.py:320-323) and Huber (the reference returns ``clamp(x - t, -1, 1)``,
nn/losses.py:391-393, the derivative of the textbook Huber loss rather than
of its own forward ``0.5 min(d^2, |d|)``).  Targets get no gradient.

There is no CPU path: every function raises ``ValueError`` for tensors that
are not on a GPU and ``RuntimeError`` when the HIP library is missing.
"""
from __future__ import annotations

import ctypes
from typing import Dict, Optional, Tuple

import numpy as np
import torch

from ._lib import ACT, BINOP, GEMM_MODE, REDOP, call, shape_array

__all__ = ["dense", "relu", "sigmoid", "softmax", "softmax_cross_entropy", "mse", "huber"]

_contexts: Dict[Tuple[int, int], int] = {}


def _ctx(t: torch.Tensor) -> int:
    """Library context bound to torch's current stream on t's device."""
    if not t.is_cuda:
        raise ValueError("vulkpy_amd.torch_autograd needs tensors on a ROCm GPU "
                         f"(got device {t.device})")
    dev = t.device.index if t.device.index is not None else torch.cuda.current_device()
    stream = torch.cuda.current_stream(dev).cuda_stream
    key = (dev, int(stream))
    h = _contexts.get(key)
    if h is None:
        out = ctypes.c_void_p()
        call("vkh_gpu_create_on_stream", dev, ctypes.c_void_p(stream), ctypes.byref(out))
        h = out.value
        _contexts[key] = h
    return h


def _f32(t: torch.Tensor, name: str) -> torch.Tensor:
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    return t.contiguous()


def _p(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _scale_by(h: int, dx: torch.Tensor, g: torch.Tensor) -> None:
    """dx *= g for a one-element device tensor g, without reading g on the host."""
    g = _f32(g.reshape(1), "grad_output")
    n = dx.numel()
    call("vkh_binary_bcast", h, BINOP["mul"], _p(dx), _p(g), _p(dx), 1,
         shape_array((n,)), shape_array((1,)), shape_array((n,)))


class _Dense(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b, precision):
        x2 = _f32(x, "x").reshape(-1, x.shape[-1])
        w = _f32(w, "w")
        if b is not None:
            b = _f32(b, "b")
        B, I = x2.shape
        O = w.shape[0]
        if w.shape != (O, I) or (b is not None and b.shape != (O,)):
            raise ValueError(f"dense: x {tuple(x.shape)}, w {tuple(w.shape)}, b "
                             f"{None if b is None else tuple(b.shape)} do not match")
        h = _ctx(x2)
        mode = GEMM_MODE[precision]
        y = torch.empty((B, O), device=x.device, dtype=torch.float32)
        call("vkh_gemm", h, mode, 0, 1, B, O, I, 1.0, _p(x2), I, _p(w), I, 0.0, _p(y), O,
             _p(b), ACT[None], None)
        ctx.save_for_backward(x2, w)
        ctx.mode, ctx.has_bias, ctx.x_shape = mode, b is not None, x.shape
        return y.reshape(*x.shape[:-1], O)

    @staticmethod
    def backward(ctx, dy):
        x2, w = ctx.saved_tensors
        B, I = x2.shape
        O = w.shape[0]
        dy = _f32(dy, "dy").reshape(B, O)
        h = _ctx(dy)
        dx = dw = db = None
        if ctx.needs_input_grad[0]:
            dx = torch.empty((B, I), device=dy.device, dtype=torch.float32)
            call("vkh_gemm", h, ctx.mode, 0, 0, B, I, O, 1.0, _p(dy), O, _p(w), I, 0.0, _p(dx), I,
                 None, ACT[None], None)
            dx = dx.reshape(ctx.x_shape)
        if ctx.needs_input_grad[1]:
            dw = torch.empty((O, I), device=dy.device, dtype=torch.float32)
            call("vkh_gemm", h, ctx.mode, 1, 0, O, I, B, 1.0, _p(dy), O, _p(x2), I, 0.0, _p(dw), I,
                 None, ACT[None], None)
        if ctx.has_bias and ctx.needs_input_grad[2]:
            db = torch.empty((O,), device=dy.device, dtype=torch.float32)
            call("vkh_bias_grad", h, _p(dy), _p(db), B, O, 0.0)
        return dx, dw, db, None


class _Act(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, act):
        x = _f32(x, "x")
        h = _ctx(x)
        y = torch.empty_like(x)
        call("vkh_activation", h, ACT[act], _p(x), _p(y), x.numel())
        ctx.save_for_backward(y)
        ctx.act = act
        return y

    @staticmethod
    def backward(ctx, dy):
        (y,) = ctx.saved_tensors
        dy = _f32(dy, "dy")
        dx = torch.empty_like(y)
        call("vkh_activation_backward", _ctx(dy), ACT[ctx.act], _p(y), _p(dy), _p(dx), y.numel())
        return dx, None


class _Softmax(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        x = _f32(x, "x")
        cols = x.shape[-1]
        rows = x.numel() // max(cols, 1)
        y = torch.empty_like(x)
        call("vkh_softmax", _ctx(x), _p(x), _p(y), rows, cols)
        ctx.save_for_backward(y)
        return y

    @staticmethod
    def backward(ctx, dy):
        (y,) = ctx.saved_tensors
        dy = _f32(dy, "dy")
        cols = y.shape[-1]
        rows = y.numel() // max(cols, 1)
        dx = torch.empty_like(y)
        call("vkh_softmax_backward", _ctx(dy), _p(y), _p(dy), _p(dx), rows, cols)
        return dx


def _reduce_rows(h: int, L: torch.Tensor, reduce: str) -> torch.Tensor:
    """Batch reduction of the per-sample losses (ReduceLoss, nn/losses.py:66-73)."""
    B = L.numel()
    out = torch.empty((), device=L.device, dtype=torch.float32)
    mul = float(np.float32(1.0 / B)) if reduce == "mean" else 1.0
    call("vkh_reduce_axis", h, REDOP["sum"], _p(L), _p(out), 1, B, 1, 0, mul, 1.0)
    return out


_KIND = {"sce": 0, "mse": 1, "huber": 3}


class _Loss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, t, kind, reduce):
        if reduce not in ("mean", "sum"):
            raise ValueError(f"reduce must be 'mean' or 'sum', got {reduce!r}")
        x = _f32(x, "x")
        t = _f32(t, "target")
        if x.dim() != 2 or t.shape != x.shape:
            raise ValueError(f"loss: x {tuple(x.shape)} and target {tuple(t.shape)} must be the same (B, C)")
        B, C = x.shape
        h = _ctx(x)
        L = torch.empty((B,), device=x.device, dtype=torch.float32)
        if kind == "sce":
            sm = torch.empty_like(x)
            call("vkh_softmax_cross_entropy", h, _p(x), _p(t), _p(sm), _p(L), B, C)
            ctx.save_for_backward(sm, t)
        elif kind == "mse":
            call("vkh_mse_rows", h, _p(x), _p(t), _p(L), B, C)
            ctx.save_for_backward(x, t)
        else:
            call("vkh_huber_rows", h, _p(x), _p(t), _p(L), B, C)
            ctx.save_for_backward(x, t)
        ctx.kind = kind
        ctx.scale = float(np.float32(1.0 / B)) if reduce == "mean" else 1.0
        return _reduce_rows(h, L, reduce)

    @staticmethod
    def backward(ctx, g):
        a, t = ctx.saved_tensors
        h = _ctx(a)
        dx = torch.empty_like(a)
        call("vkh_loss_grad", h, _KIND[ctx.kind], _p(a), _p(t), _p(dx), a.numel(), ctx.scale)
        _scale_by(h, dx, g)
        return dx, None, None, None


def dense(x: torch.Tensor, w: torch.Tensor, b: Optional[torch.Tensor] = None, *,
          precision: str = "fp32") -> torch.Tensor:
    """``x @ w.T + b`` with w of shape (out, in), as vulkpy's Dense.
    ``precision`` is the product precision ("fp32", "bf16" or "bf16x3"; FP32
    accumulation in every mode) of the forward and of both backward products."""
    if precision not in GEMM_MODE:
        raise ValueError(f"unknown precision {precision!r}")
    return _Dense.apply(x, w, b, precision)


def relu(x: torch.Tensor) -> torch.Tensor:
    """max(x, 0) (numpy.maximum semantics; NaN propagates)."""
    return _Act.apply(x, "relu")


def sigmoid(x: torch.Tensor) -> torch.Tensor:
    """1 / (1 + exp(-x))."""
    return _Act.apply(x, "sigmoid")


def softmax(x: torch.Tensor) -> torch.Tensor:
    """Softmax over the last axis."""
    return _Softmax.apply(x)


def softmax_cross_entropy(x: torch.Tensor, target: torch.Tensor, reduce: str = "mean") -> torch.Tensor:
    """``-sum(t * log(softmax(x) + 1e-8))`` per row, reduced over the batch."""
    return _Loss.apply(x, target, "sce", reduce)


def mse(x: torch.Tensor, target: torch.Tensor, reduce: str = "mean") -> torch.Tensor:
    """``sum((t - x)^2)`` per row, reduced over the batch."""
    return _Loss.apply(x, target, "mse", reduce)


def huber(x: torch.Tensor, target: torch.Tensor, reduce: str = "mean") -> torch.Tensor:
    """``sum(0.5 min(d^2, |d|))`` with ``d = t - x`` per row, reduced over the batch."""
    return _Loss.apply(x, target, "huber", reduce)
