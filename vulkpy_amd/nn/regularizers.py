"""
Neural Network Regularizer Module (:mod:`vulkpy_amd.nn.regularizers`)
=====================================================================

Same as vulkpy/nn/regularizers.py:23-192.
"""
from __future__ import annotations
import logging

from vulkpy_amd.vkarray import Array
from .core import Regularizer

__all__ = [
    "Lasso",
    "Ridge",
    "Elastic",
]

logger = logging.getLogger("vulkpy")


class Lasso(Regularizer):
    r"""
    Lasso (L1) Regularization: L = coeff * sum |W|, dL/dW = coeff * sign(W)
    """
    def __init__(self, coeff: float = 1.0):
        logger.debug(f"Lasso(L1={coeff})")
        self.coeff: float = coeff

    def loss(self, param: Array) -> Array:
        L = param.abs().sum()
        L *= self.coeff
        return L

    def grad(self, param: Array) -> Array:
        return self.coeff * param.sign()


class Ridge(Regularizer):
    r"""
    Ridge (L2) Regularization: L = coeff * sum W^2, dL/dW = 2 coeff W
    """
    def __init__(self, coeff: float = 1.0):
        logger.debug(f"Ridge(L2={coeff})")
        self.coeff: float = coeff

    def loss(self, param: Array) -> Array:
        L = (param ** 2).sum()
        L *= self.coeff
        return L

    def grad(self, param: Array) -> Array:
        return (2 * self.coeff) * param


class Elastic(Regularizer):
    r"""
    Elastic (L1 + L2) Regularization
    """
    def __init__(self, L1: float = 1.0, L2: float = 1.0):
        self.L1 = Lasso(L1)
        self.L2 = Ridge(L2)

    def loss(self, param: Array) -> Array:
        return self.L1.loss(param) + self.L2.loss(param)

    def grad(self, param: Array) -> Array:
        return self.L1.grad(param) + self.L2.grad(param)
