"""
Neural Network Loss Module (:mod:`vulkpy_amd.nn.losses`)
========================================================

Same API as vulkpy/nn/losses.py:40-458.  For 2-D inputs the per-sample loss
rows and the gradients are single fused kernels performing the reference's
FP32 operations in its order.

Examples
--------
>>> import vulkpy_amd as vk
>>> from vulkpy_amd import nn
>>> L = nn.CrossEntropyLoss()
>>> loss = L(x, y)
>>> dx = L.grad()
"""
from __future__ import annotations
from typing import cast, Callable, Iterable, Literal, Optional, Tuple

from vulkpy_amd._lib import call
from vulkpy_amd.vkarray import Array
from .core import Loss
from .layers import Softmax

__all__ = [
    "CrossEntropyLoss",
    "SoftmaxCrossEntropyLoss",
    "MSELoss",
    "HuberLoss",
    "MixLoss",
]

# loss gradient kinds of vkh_loss_grad
_SCE, _MSE, _HUBER = 0, 1, 2


F = Callable[[Array], Array]
class ReduceLoss(Loss):
    def __init__(self, reduce: Literal["mean", "sum"] = "mean"):
        tmp: Tuple[F, Optional[F]] = {
            "mean": (lambda _L: _L.mean(axis=0), lambda _dx: 1/_dx.shape[0]),
            "sum": (lambda _L: _L.sum(axis=0), None),
        }[reduce]
        self.reduce, self.scale_backward = tmp

    def __call__(self, x: Array, y: Array) -> Array:
        """
        Compute Loss (per-sample ``forward`` reduced over the batch)
        """
        self._x = x
        self._y = y
        L = self.forward(x, y)
        return self.reduce(L)

    def grad(self) -> Array:
        """
        Compute Gradients of the last ``__call__(x, y)`` (scaled by 1/B for
        ``reduce="mean"``, vulkpy/nn/losses.py:70-86)
        """
        fused = getattr(type(self), "_backward_scaled", None)
        if fused is not None and _backward_not_overridden(self):
            scale = 1.0
            if self.scale_backward is not None:
                scale = self.scale_backward(self._x)
            return self._backward_scaled(float(scale))
        dx = self.backward()
        if self.scale_backward is not None:
            dx *= self.scale_backward(dx)
        return dx

    def forward(self, x: Array, y: Array) -> Array:
        raise NotImplementedError

    def backward(self) -> Array:
        raise NotImplementedError


def _backward_not_overridden(loss) -> bool:
    for klass in type(loss).__mro__:
        if "_backward_scaled" in vars(klass):
            return type(loss).backward is klass.backward
    return False


def _rows(x: Array, y: Array) -> Tuple[int, int]:
    if tuple(x.shape) != tuple(y.shape):
        raise ValueError(f"Incompatible shapes: {x.shape} vs {y.shape}")
    return int(x.shape[0]), int(x.shape[1])


def _grad_kernel(x: Array, y: Array, kind: int, scale: float) -> Array:
    if tuple(x.shape) != tuple(y.shape):
        raise ValueError(f"Incompatible shapes: {x.shape} vs {y.shape}")
    dx = Array(x._gpu, shape=x.shape)
    call("vkh_loss_grad", x._gpu._handle, kind, x._ptr, y._ptr, dx._ptr, x.size, float(scale))
    dx._set_job([x, y])
    return dx


class CrossEntropyLoss(ReduceLoss):
    """
    Cross Entropy Loss
    """
    def __init__(self, *args, **kwargs):
        """
        Parameters
        ----------
        reduce : {"mean", "sum"}, optional
            Reduction method over batch. The default is ``"mean"``.
        """
        super().__init__(*args, **kwargs)

    def forward(self, x: Array, y: Array) -> Array:
        r"""
        Forward: :math:`L = - \sum_i y_i \log (x_i + 10^{-8})` per sample
        (nn_cross_entropy.comp:26 then sum(axis=1))
        """
        if len(x.shape) == 2:
            B, C = _rows(x, y)
            L = Array(x._gpu, shape=(B,))
            call("vkh_cross_entropy", x._gpu._handle, x._ptr, y._ptr, L._ptr, B, C)
            L._set_job([x, y])
            return L
        L = (x + 1e-8).log()
        L *= (0.0 - y)
        return L.sum(axis=1)

    def _backward_scaled(self, scale: float) -> Array:
        if tuple(self._x.shape) != tuple(self._y.shape):
            raise ValueError(f"Incompatible shapes: {self._x.shape} vs {self._y.shape}")
        dx = Array(self._x._gpu, shape=self._x.shape)
        call("vkh_cross_entropy_backward", self._x._gpu._handle, self._x._ptr, self._y._ptr,
             dx._ptr, self._x.size, scale)
        dx._set_job([self._x, self._y])
        return dx

    def backward(self) -> Array:
        r"""
        Backward: :math:`dx = -y / (x + 10^{-8})` (nn_cross_entropy_backward.comp)
        """
        return self._backward_scaled(1.0)


class SoftmaxCrossEntropyLoss(CrossEntropyLoss):
    """
    Softmax Cross Entropy Loss
    """
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._sm = Softmax()

    def forward(self, x: Array, y: Array) -> Array:
        r"""
        Forward: :math:`L = - \sum_i y_i \log (\rm{softmax}(x)_i + 10^{-8})`
        (softmax and cross entropy fused per row)
        """
        if len(x.shape) == 2:
            B, C = _rows(x, y)
            sm = Array(x._gpu, shape=x.shape)
            L = Array(x._gpu, shape=(B,))
            call("vkh_softmax_cross_entropy", x._gpu._handle, x._ptr, y._ptr, sm._ptr, L._ptr, B, C)
            sm._set_job([x])
            L._set_job([x, y, sm])
            self._sm._x = x
            self._sm._y = sm
            return L
        return super().forward(self._sm(x), y)

    def _backward_scaled(self, scale: float) -> Array:
        return _grad_kernel(cast(Array, self._sm._y), self._y, _SCE, scale)

    def backward(self) -> Array:
        r"""
        Backward: :math:`dx = \rm{softmax}(x) - y` (nn/losses.py:249)
        """
        return self._backward_scaled(1.0)


class MSELoss(ReduceLoss):
    """
    Mean Squared Loss
    """
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)

    def forward(self, x: Array, y: Array) -> Array:
        r"""
        Forward: :math:`L = \sum_i (y_i - x_i)^2` per sample (nn/losses.py:294-296)
        """
        if len(x.shape) == 2:
            B, C = _rows(x, y)
            L = Array(x._gpu, shape=(B,))
            call("vkh_mse_rows", x._gpu._handle, x._ptr, y._ptr, L._ptr, B, C)
            L._set_job([x, y])
            return L
        L = (y - x)
        L **= 2.0
        return L.sum(axis=1)

    def _backward_scaled(self, scale: float) -> Array:
        return _grad_kernel(self._x, self._y, _MSE, scale)

    def backward(self) -> Array:
        r"""
        Backward: :math:`dx = 2 (x - y)` (nn/losses.py:318-320)
        """
        return self._backward_scaled(1.0)


class HuberLoss(ReduceLoss):
    """
    Huber Loss
    """
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)

    def forward(self, x: Array, y: Array) -> Array:
        r"""
        Forward: :math:`L = 0.5 \sum_i \min(|x_i - y_i|^2, |x_i - y_i|)` per
        sample (nn/losses.py:365-369)
        """
        if len(x.shape) == 2:
            B, C = _rows(x, y)
            L = Array(x._gpu, shape=(B,))
            call("vkh_huber_rows", x._gpu._handle, x._ptr, y._ptr, L._ptr, B, C)
            L._set_job([x, y])
            return L
        delta = y - x
        delta.abs(inplace=True)
        delta.min(delta ** 2.0, inplace=True)
        delta *= 0.5
        return delta.sum(axis=1)

    def _backward_scaled(self, scale: float) -> Array:
        return _grad_kernel(self._x, self._y, _HUBER, scale)

    def backward(self) -> Array:
        r"""
        Backward: :math:`dx = \text{clamp}(x - y, -1.0, 1.0)` (nn/losses.py:391-393)
        """
        return self._backward_scaled(1.0)


class MixLoss(Loss):
    """
    Mixing Loss class
    """
    def __init__(self, losses: Iterable[Tuple[float, Loss]]):
        """
        Parameters
        ----------
        losses : iterable of tuple of float and vulkpy_amd.nn.Loss
            Sets of coefficient and loss.

        Raises
        ------
        ValueError
            When losses is empty
        """
        self.L: Tuple[Tuple[float, Loss], ...] = tuple(losses)
        if len(self.L) < 1:
            raise ValueError(f"losses should not empty")

    def __call__(self, x: Array, y: Array) -> Array:
        return self._sum(lambda _L: _L(x, y))

    def grad(self) -> Array:
        return self._sum(lambda _L: _L.grad())

    def _sum(self, F: Callable[[Loss], Array]) -> Array:
        coeff, _L = self.L[0]
        s = coeff * F(_L)

        for coeff, _L in self.L[1:]:
            s += coeff * F(_L)

        return s
