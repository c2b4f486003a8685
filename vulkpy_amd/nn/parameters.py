"""
Neural Network Parameter (:mod:`vulkpy_amd.nn.parameters`)
==========================================================

Same as vulkpy/nn/parameters.py:14-116.
"""
from __future__ import annotations
from typing import Callable, Iterable, Optional

from vulkpy_amd.vkarray import GPU, Array, zeros
from .core import Optimizer, OptimizerState, Regularizer
from .optimizers import Adam


__all__ = [
    "Parameter"
]


def _has_fused_update(state) -> bool:
    """True when the state's class provides ``_update`` and does not override
    the ``grad2diff`` that the fused kernel reproduces."""
    for klass in type(state).__mro__:
        if "_update" in vars(klass):
            return type(state).grad2diff is klass.grad2diff
    return False


class Parameter:
    """
    Neural Network Parameter
    """
    def __init__(self,
                 gpu: GPU,
                 shape: Iterable[int],
                 trainable: bool = True,
                 opt: Optional[Optimizer] = None,
                 initializer: Optional[Callable[[GPU, Iterable[int]], Array]]=None,
                 regularizer: Optional[Regularizer] = None):
        if initializer is None:
            initializer = zeros
        self.value: Array = initializer(gpu, shape)

        self.grad: Optional[Array] = None
        self.opt_state: Optional[OptimizerState] = None
        if trainable:
            self.grad = zeros(gpu, shape=shape)

            if opt is None:
                opt = Adam(gpu)
            self.opt_state = opt.init_state(shape)

        self.R: Optional[Regularizer] = regularizer

    def is_trainable(self) -> bool:
        return self.grad is not None

    def add_grad(self, grad: Array):
        if self.grad is not None:
            self.grad += grad

    def zero_grad(self):
        if self.grad is not None:
            self.grad[:] = 0.0

    def update(self):
        """
        Update value with accumulated gradients only if this value is trainable.
        """
        if self.grad is not None:
            if _has_fused_update(self.opt_state):
                self.opt_state._update(self.value, self.grad)
            else:
                self.value += self.opt_state.grad2diff(self.grad)

    def regular_loss(self) -> Array:
        if self.R is not None:
            return self.R.loss(self.value)

        return zeros(self.value._gpu, shape=(1,))

    def regular_grad(self):
        if self.R is not None:
            self.add_grad(self.R.grad(self.value))
