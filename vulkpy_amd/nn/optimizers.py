"""
Neural Network Optimizer Module (:mod:`vulkpy_amd.nn.optimizers`)
=================================================================

Same API as vulkpy/nn/optimizers.py:24-350.  ``grad2diff`` keeps the
reference's Array-op composition; ``Parameter.update`` uses the fused
``_update`` kernels, which apply the same FP32 operations in the same order
(so both paths give identical weights).
"""
from __future__ import annotations
import logging
from typing import Iterable

from vulkpy_amd._lib import call
from vulkpy_amd.vkarray import GPU, Array, zeros
from .core import Optimizer, OptimizerState

__all__ = [
    "SGD", "SGDState",
    "AdaGrad", "AdaGradState",
    "Adam", "AdamState",
    "Optimizer", "OptimizerState",
]

logger = logging.getLogger("vulkpy")


class SGDState(OptimizerState):
    """
    Optimizer State for SGD
    """
    def __init__(self, opt: SGD):
        self.opt: SGD = opt

    def grad2diff(self, grad: Array) -> Array:
        """diff = (-lr) * grad (vulkpy/nn/optimizers.py:53)"""
        return (-self.opt.lr) * grad

    def _update(self, value: Array, grad: Array):
        call("vkh_sgd", value._gpu._handle, value._ptr, grad._ptr, value.size, float(self.opt.lr))
        value._set_job([grad])


class SGD(Optimizer):
    """
    SGD Optimizer (constant learning rate)
    """
    def __init__(self, lr: float):
        self.lr: float = lr
        logger.debug("SGD(lr=%f)", self.lr)

    def init_state(self, shape: Iterable[int]) -> SGDState:
        return SGDState(self)


class AdaGradState(OptimizerState):
    """
    Optimizer State for AdaGrad
    """
    def __init__(self, opt: AdaGrad, shape: Iterable[int], tau: float):
        self.opt: AdaGrad = opt
        self.h: Array = zeros(self.opt.gpu, shape=shape)
        self.h[:] = tau

    def grad2diff(self, grad: Array) -> Array:
        """vulkpy/nn/optimizers.py:134-140"""
        self.h += (grad ** 2)

        sqrt = self.h.sqrt()  #               sqrt(sum)
        sqrt += self.opt.eps  #               sqrt(sum) + eps
        ret = grad / sqrt     #       grad / (sqrt(sum) + eps)
        ret *= (-self.opt.lr) # -lr * grad / (sqrt(sum) + eps)
        return ret

    def _update(self, value: Array, grad: Array):
        call("vkh_adagrad", value._gpu._handle, value._ptr, grad._ptr, self.h._ptr, value.size,
             float(self.opt.lr), float(self.opt.eps))
        value._set_job([grad])
        self.h._set_job()


class AdaGrad(Optimizer):
    """
    AdaGrad Optimizer
    """
    def __init__(self,
                 gpu: GPU, *,
                 lr: float = 0.01,
                 tau: float = 0.0,
                 eps: float = 1e-8):
        self.gpu: GPU = gpu
        self.lr: float = lr
        self.tau: float = tau
        self.eps: float = eps

        logger.debug("AdaGrad(lr=%f, tau=%f, eps=%f)",
                     self.lr, self.tau, self.eps)

    def init_state(self, shape: Iterable[int]) -> AdaGradState:
        return AdaGradState(opt=self, shape=shape, tau=self.tau)


class AdamState(OptimizerState):
    """
    Optimizer State for Adam
    """
    def __init__(self, opt: Adam, shape: Iterable[int]):
        self.opt: Adam = opt
        self.m: Array = zeros(self.opt.gpu, shape=shape)
        self.v: Array = zeros(self.opt.gpu, shape=shape)
        self.beta1t: float = 1.0
        self.beta2t: float = 1.0

    def grad2diff(self, grad: Array) -> Array:
        """vulkpy/nn/optimizers.py:235-253"""
        self.m *= self.opt.beta1
        self.m += (1 - self.opt.beta1) * grad        # Allocate

        self.v *= self.opt.beta2
        self.v += (1 - self.opt.beta2) * (grad ** 2) # Allocate

        self.beta1t *= self.opt.beta1
        self.beta2t *= self.opt.beta2

        mhat = self.m / (1 - self.beta1t) # Allocate
        vhat = self.v / (1 - self.beta2t) # Allocate

        vhat.sqrt(inplace=True) # sqrt(vhat)
        vhat += self.opt.eps    # sqrt(vhat) + eps

        mhat *= (-self.opt.lr)  # -lr * mhat
        mhat /= vhat            # -lr * mhat / (sqrt(vhat) + eps)

        return mhat

    def _update(self, value: Array, grad: Array):
        """Fused grad2diff + ``value += diff`` (one kernel)."""
        self.beta1t *= self.opt.beta1
        self.beta2t *= self.opt.beta2
        call("vkh_adam", value._gpu._handle, value._ptr, grad._ptr, self.m._ptr, self.v._ptr,
             value.size, float(self.opt.lr), float(self.opt.beta1), float(self.opt.beta2),
             float(self.opt.eps), float(self.beta1t), float(self.beta2t))
        value._set_job([grad])
        self.m._set_job()
        self.v._set_job()


class Adam(Optimizer):
    r"""
    Adam Optimizer

    .. math::

         m_t = \beta _1 m_{t-1} + (1 - \beta _1) g_t\\
         v_t = \beta _2 v_{t-1} + (1 - \beta _2) g_t ^2\\
         \theta _t = \theta _{t-1} - \text{lr} \times
         \hat{m}_t/(\sqrt{\hat{v}_t} + \epsilon)

    References
    ----------
    .. [adam1] D. Kingma and J. Ba, "Adam: A Method for Stochastic Optimization",
       ICLR (Poster) 2015
    """
    def __init__(self,
                 gpu: GPU, *,
                 lr: float = 0.001,
                 beta1: float = 0.9,
                 beta2: float = 0.999,
                 eps: float = 1e-8):
        self.gpu: GPU = gpu
        self.lr: float = lr
        self.beta1: float = beta1
        self.beta2: float = beta2
        self.eps: float = eps

        logger.debug("Adam(lr=%f, beta1=%f, beta2=%f, eps=%f)",
                     self.lr, self.beta1, self.beta2, self.eps)

    def init_state(self, shape: Iterable[int]) -> AdamState:
        return AdamState(opt=self, shape=shape)
