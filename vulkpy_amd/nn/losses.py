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


def _inverse_batch(dx: Array) -> float:
    """gradient scale of a batch mean: 1 / B (vulkpy/nn/losses.py:43)"""
    return 1 / dx.shape[0]


class ReduceLoss(Loss):
    """
    A loss evaluated per sample by ``forward`` and reduced over the batch
    (axis 0) by its mean or its sum (vulkpy/nn/losses.py:40-86).

    Parameters
    ----------
    reduce : {"mean", "sum"}, optional
        Batch reduction (default ``"mean"``).  The gradient of the mean is
        the per-sample gradient scaled by 1 / B.
    """
    def __init__(self, reduce: Literal["mean", "sum"] = "mean"):
        if reduce not in ("mean", "sum"):
            raise KeyError(reduce)
        self._batch_reduction = reduce

    def reduce(self, L: Array) -> Array:
        """batch reduction of the per-sample losses ``L``"""
        if self._batch_reduction == "mean":
            return L.mean(axis=0)
        return L.sum(axis=0)

    @property
    def scale_backward(self) -> Optional[Callable[[Array], float]]:
        """scale of the gradient as a function of the gradient array
        (1 / B for the mean), or ``None`` (the sum)"""
        return _inverse_batch if self._batch_reduction == "mean" else None

    def __call__(self, x: Array, y: Array) -> Array:
        """
        Loss of the batch ``(x, y)``; ``grad()`` then differentiates it
        """
        self._x = x
        self._y = y
        return self.reduce(self.forward(x, y))

    def grad(self) -> Array:
        """
        Gradient of the last ``__call__(x, y)`` with respect to ``x``
        """
        scale = self.scale_backward
        if getattr(type(self), "_backward_scaled", None) is not None and _backward_not_overridden(self):
            # one kernel: the per-sample gradient with the batch scale applied
            return self._backward_scaled(1.0 if scale is None else float(scale(self._x)))
        dx = self.backward()
        if scale is not None:
            dx *= scale(dx)
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
    Weighted sum of losses, sum_k c_k L_k (vulkpy/nn/losses.py:396-458); its
    gradient is the same weighted sum of the losses' gradients.  Terms are
    scaled and accumulated left to right in FP32.

    Parameters
    ----------
    losses : iterable of (float, vulkpy_amd.nn.Loss)
        Coefficient and loss of every term.

    Raises
    ------
    ValueError
        When ``losses`` is empty
    """
    def __init__(self, losses: Iterable[Tuple[float, Loss]]):
        self.L: Tuple[Tuple[float, Loss], ...] = tuple(losses)
        if not self.L:
            raise ValueError(f"losses should not empty")

    def __call__(self, x: Array, y: Array) -> Array:
        """weighted sum of every loss of the batch ``(x, y)``"""
        return self._weighted(lambda loss: loss(x, y))

    def grad(self) -> Array:
        """weighted sum of every loss's gradient for the last call"""
        return self._weighted(lambda loss: loss.grad())

    def _weighted(self, term: Callable[[Loss], Array]) -> Array:
        total: Optional[Array] = None
        for coeff, loss in self.L:
            scaled = coeff * term(loss)
            if total is None:
                total = scaled
            else:
                total += scaled
        return total
