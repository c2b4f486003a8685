"""
Neural Network Core Module (:mod:`vulkpy_amd.nn.core`)
======================================================

Abstract base classes (same as vulkpy/nn/core.py:22-337).
"""
from __future__ import annotations
from typing import Iterable

from vulkpy_amd.vkarray import GPU, Array


__all__ = [
    "OptimizerState",
    "Optimizer",
    "Regularizer",
    "Loss",
    "Module",
]


class OptimizerState:
    """
    Abstract base class for Optimizer State

    Subclass of ``OptimizerState`` should implement ``grad2diff()``,
    which takes accumulated gradients and returns update difference.
    """
    def grad2diff(self, grad: Array) -> Array:
        """
        Compute update diff from gradient (``v += opt_state.grad2diff(grad)``)
        """
        raise NotImplementedError


class Optimizer:
    """
    Abstract base class for Optimizer

    ``Optimizer.init_state()`` returns the per-parameter ``OptimizerState``.
    """
    def init_state(self, shape: Iterable[int]) -> OptimizerState:
        raise NotImplementedError


class Loss:
    """
    Abstract base class for Loss

    Subclass of ``Loss`` must implements ``__call__()`` and ``grad()``.
    """
    def __call__(self, x: Array, y: Array) -> Array:
        raise NotImplementedError

    def grad(self) -> Array:
        raise NotImplementedError


class Regularizer:
    """
    Abstract base class for Regularizer

    Subclass must implement ``loss()`` and ``grad()``.
    """
    def loss(self, param: Array) -> Array:
        raise NotImplementedError

    def grad(self, param: Array) -> Array:
        raise NotImplementedError


class Module:
    """
    Abstract base class for Module

    Subclass must implement ``forward()`` and ``backward()``, and can implement
    ``zero_grad()`` and ``update()`` when it is necessary.
    """

    def __call__(self, x: Array) -> Array:
        """
        Call Module

        Raises
        ------
        ValueError
            If input (``x``) shape doesn't have at least 2-dimensions.

        Notes
        -----
        This function stores input (``x``) and output (``y``) for training.
        """
        if len(x.shape) < 2:
            raise ValueError("Input must have at least 2-dimensions.")

        self._x = x
        self._y = self.forward(x)
        return self._y

    def forward(self, x: Array) -> Array:
        raise NotImplementedError

    def backward(self, dy: Array) -> Array:
        raise NotImplementedError

    def zero_grad(self):
        pass

    def update(self):
        pass
