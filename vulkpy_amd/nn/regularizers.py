"""
Neural Network Regularizer Module (:mod:`vulkpy_amd.nn.regularizers`)
=====================================================================

Lasso (L1), Ridge (L2) and Elastic (L1 + L2) penalties with the interface of
vulkpy/nn/regularizers.py:23-192.

Each regularizer is described by two coefficients, an L1 weight ``l1`` and
an L2 weight ``l2`` (Lasso: ``(coeff, 0)``, Ridge: ``(0, coeff)``, Elastic:
``(L1, L2)``), and its gradient is one fused kernel (``vkh_regularizer_grad``)
that forms the reference's terms in its FP32 order: ``coeff * sign(W)``,
``(2 coeff) * W`` with the scalar ``2 coeff`` rounded once, and for Elastic
the sum of the two rounded terms.  ``Parameter.regular_grad`` adds that
term onto the gradient inside the same kernel, and a captured training step
folds it into the optimizer's update kernel.
"""
from __future__ import annotations
import logging
from typing import Tuple

import numpy as np

from vulkpy_amd._lib import call
from vulkpy_amd.vkarray import Array
from .core import Regularizer

__all__ = [
    "Lasso",
    "Ridge",
    "Elastic",
]

logger = logging.getLogger("vulkpy")

# vkh_regularizer_grad kinds
_L1, _L2, _L1L2 = 1, 2, 3


class _Penalty(Regularizer):
    """Common base: the fused gradient kernel for the ``terms()`` of a penalty."""

    _kind = 0

    def terms(self) -> Tuple[int, float, float]:
        """(kernel kind, l1 coefficient, FP32 scalar 2 * l2) of the gradient"""
        raise NotImplementedError

    def _launch(self, param: Array, grad: Array, accumulate: bool):
        kind, l1, s2 = self.terms()
        call("vkh_regularizer_grad", param._gpu._handle, kind, param._ptr, grad._ptr, param.size,
             l1, s2, int(accumulate))

    def grad(self, param: Array) -> Array:
        out = Array(param._gpu, shape=param.shape)
        self._launch(param, out, accumulate=False)
        out._set_job([param])
        return out

    def _add_grad(self, param: Array, grad: Array):
        """grad += self.grad(param), in one kernel"""
        self._launch(param, grad, accumulate=True)
        grad._set_job([param])


def _scalar(x: float) -> float:
    """a Python scalar as the FP32 value a scalar kernel receives"""
    return float(np.float32(x))


class Lasso(_Penalty):
    r"""
    Lasso (L1) Regularization

    .. math::

         L = \text{coeff} \sum_i |W_i|, \qquad
         \partial L / \partial W_i = \text{coeff} \cdot \mathrm{sign}(W_i)
    """
    def __init__(self, coeff: float = 1.0):
        """
        Parameters
        ----------
        coeff : float, optional
            L1 coefficient
        """
        logger.debug(f"Lasso(L1={coeff})")
        self.coeff: float = coeff

    def terms(self) -> Tuple[int, float, float]:
        return _L1, _scalar(self.coeff), 0.0

    def loss(self, param: Array) -> Array:
        """coeff * sum |W| (an FP64 sum rounded once, then the FP32 scale)"""
        total = param.abs().sum()
        total *= self.coeff
        return total


class Ridge(_Penalty):
    r"""
    Ridge (L2) Regularization

    .. math::

         L = \text{coeff} \sum_i W_i^2, \qquad
         \partial L / \partial W_i = 2\,\text{coeff} \cdot W_i
    """
    def __init__(self, coeff: float = 1.0):
        """
        Parameters
        ----------
        coeff : float, optional
            L2 coefficient
        """
        logger.debug(f"Ridge(L2={coeff})")
        self.coeff: float = coeff

    def terms(self) -> Tuple[int, float, float]:
        return _L2, 0.0, _scalar(2 * self.coeff)

    def loss(self, param: Array) -> Array:
        """coeff * sum W^2"""
        total = (param ** 2).sum()
        total *= self.coeff
        return total


class Elastic(_Penalty):
    r"""
    Elastic (L1 + L2) Regularization

    .. math::

         L = \alpha \sum_i |W_i| + \beta \sum_i W_i^2, \qquad
         \partial L / \partial W_i = \alpha\,\mathrm{sign}(W_i) + 2 \beta W_i
    """
    def __init__(self, L1: float = 1.0, L2: float = 1.0):
        """
        Parameters
        ----------
        L1 : float, optional
            L1 coefficient (alpha)
        L2 : float, optional
            L2 coefficient (beta)
        """
        self.L1 = Lasso(L1)
        self.L2 = Ridge(L2)

    def terms(self) -> Tuple[int, float, float]:
        return _L1L2, _scalar(self.L1.coeff), _scalar(2 * self.L2.coeff)

    def loss(self, param: Array) -> Array:
        """the Lasso part plus the Ridge part, each rounded"""
        return self.L1.loss(param) + self.L2.loss(param)


def _fused_penalty(R) -> bool:
    """True when R's gradient is exactly the fused kernel's (a penalty of this
    module whose ``grad`` no subclass replaced)."""
    return isinstance(R, _Penalty) and type(R).grad is _Penalty.grad
