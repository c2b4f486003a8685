"""
Neural Network Parameter (:mod:`vulkpy_amd.nn.parameters`)
==========================================================

``Parameter`` with the interface of vulkpy/nn/parameters.py:14-116: a value,
an accumulated gradient (absent for frozen parameters), an optimizer state and
an optional regularizer.

On this backend every method is stream ordered and none waits on the host:

* ``zero_grad`` enqueues a fill kernel (the reference writes the mapped
  memory from the host, parameters.py:86);
* ``update`` runs the optimizer state's fused update kernel when the state
  class provides one (same FP32 operations as ``grad2diff`` + ``+=``);
* ``regular_grad`` adds the regularizer's gradient onto ``grad`` in one
  kernel for the penalties of :mod:`vulkpy_amd.nn.regularizers`.
"""
from __future__ import annotations
from typing import Callable, Iterable, Optional

from vulkpy_amd.vkarray import GPU, Array, zeros
from .core import Optimizer, OptimizerState, Regularizer
from .optimizers import Adam
from .regularizers import _fused_penalty


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
    Parameter of a layer: value, gradient, optimizer state and regularizer

    Parameters
    ----------
    gpu : vulkpy_amd.GPU
        Device holding the value, gradient and optimizer state.
    shape : iterable of int
        Shape of the value.
    trainable : bool, optional
        Keep a gradient and an optimizer state (default ``True``).
    opt : vulkpy_amd.nn.Optimizer, optional
        Optimizer; ``Adam(gpu)`` when ``None``.
    initializer : callable, optional
        ``initializer(gpu, shape)`` returning the initial value; zeros when
        ``None``.
    regularizer : vulkpy_amd.nn.Regularizer, optional
        Penalty whose gradient ``regular_grad`` adds.
    """
    def __init__(self,
                 gpu: GPU,
                 shape: Iterable[int],
                 trainable: bool = True,
                 opt: Optional[Optimizer] = None,
                 initializer: Optional[Callable[[GPU, Iterable[int]], Array]]=None,
                 regularizer: Optional[Regularizer] = None):
        make_value = zeros if initializer is None else initializer
        self.value: Array = make_value(gpu, shape)
        self.R: Optional[Regularizer] = regularizer
        self.grad: Optional[Array] = None
        self.opt_state: Optional[OptimizerState] = None
        if not trainable:
            return
        self.grad = zeros(gpu, shape=shape)
        optimizer = Adam(gpu) if opt is None else opt
        self.opt_state = optimizer.init_state(shape)

    def is_trainable(self) -> bool:
        """Whether the parameter keeps a gradient"""
        return self.grad is not None

    def add_grad(self, grad: Array):
        """Accumulate ``grad`` (ignored for a frozen parameter)"""
        if not self.is_trainable():
            return
        self.grad += grad

    def zero_grad(self):
        """Reset the gradient to 0 (a fill kernel after earlier work on it)"""
        if not self.is_trainable():
            return
        self.grad._fill(0.0)

    def update(self):
        """Apply the optimizer step to ``value`` (trainable parameters only)"""
        if not self.is_trainable():
            return
        state = self.opt_state
        if _has_fused_update(state):
            state._update(self.value, self.grad)
            return
        step = state.grad2diff(self.grad)
        self.value += step

    def regular_loss(self) -> Array:
        """Regularization loss of the value; a (1,) zero without a regularizer"""
        if self.R is None:
            return zeros(self.value._gpu, shape=(1,))
        return self.R.loss(self.value)

    def regular_grad(self):
        """Add the regularizer's gradient to ``grad``"""
        if self.R is None or not self.is_trainable():
            return
        if _fused_penalty(self.R):
            self.R._add_grad(self.value, self.grad)
            return
        self.add_grad(self.R.grad(self.value))
