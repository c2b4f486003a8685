"""
Neural Network Optimizer Module (:mod:`vulkpy_amd.nn.optimizers`)
=================================================================

SGD, AdaGrad and Adam with the interface of vulkpy/nn/optimizers.py:24-350.

Each state has two ways to step a parameter:

* ``grad2diff(grad)`` returns the update as an Array built from array ops,
  in the FP32 operation order of the reference's methods (cited on each), for
  callers that apply the difference themselves;
* ``_update(value, grad)``, used by ``Parameter.update``, is one fused kernel
  that performs those same FP32 operations in the same order and adds the
  result to ``value``, so both ways leave identical weights.
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

_NO_REG = (0, 0.0, 0.0)   # regularizer terms of the update kernels: none


class SGDState(OptimizerState):
    """
    State of SGD for one parameter (SGD keeps nothing but its optimizer)
    """
    def __init__(self, opt: SGD):
        self.opt: SGD = opt

    def grad2diff(self, grad: Array) -> Array:
        """(-lr) * grad (vulkpy/nn/optimizers.py:53)"""
        step = (-self.opt.lr) * grad
        return step

    def _update(self, value: Array, grad: Array):
        call("vkh_sgd", value._gpu._handle, value._ptr, grad._ptr, value.size, float(self.opt.lr),
             *_NO_REG)
        value._set_job([grad])


class SGD(Optimizer):
    """
    Stochastic gradient descent with a constant learning rate
    """
    def __init__(self, lr: float):
        self.lr: float = lr
        logger.debug("SGD(lr=%f)", self.lr)

    def init_state(self, shape: Iterable[int]) -> SGDState:
        return SGDState(self)


class AdaGradState(OptimizerState):
    """
    State of AdaGrad for one parameter: the running sum ``h`` of squared
    gradients, starting at ``tau``
    """
    def __init__(self, opt: AdaGrad, shape: Iterable[int], tau: float):
        self.opt: AdaGrad = opt
        self.h: Array = zeros(self.opt.gpu, shape=shape)
        self.h[:] = tau

    def grad2diff(self, grad: Array) -> Array:
        """h += g^2; -lr * g / (sqrt(h) + eps), in the order of
        vulkpy/nn/optimizers.py:134-140"""
        self.h += (grad ** 2)
        denom = self.h.sqrt()
        denom += self.opt.eps
        step = grad / denom
        step *= (-self.opt.lr)
        return step

    def _update(self, value: Array, grad: Array):
        call("vkh_adagrad", value._gpu._handle, value._ptr, grad._ptr, self.h._ptr, value.size,
             float(self.opt.lr), float(self.opt.eps), *_NO_REG)
        value._set_job([grad])
        self.h._set_job()


class AdaGrad(Optimizer):
    """
    AdaGrad Optimizer (per-element learning rate lr / (sqrt(sum g^2) + eps))
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
    State of Adam for one parameter: first and second moments ``m``, ``v``
    and the powers ``beta1t``, ``beta2t`` of the decay rates (Python floats,
    like the reference's)
    """
    def __init__(self, opt: Adam, shape: Iterable[int]):
        self.opt: Adam = opt
        self.m: Array = zeros(self.opt.gpu, shape=shape)
        self.v: Array = zeros(self.opt.gpu, shape=shape)
        self.beta1t: float = 1.0
        self.beta2t: float = 1.0

    def _advance_powers(self):
        self.beta1t *= self.opt.beta1
        self.beta2t *= self.opt.beta2

    def grad2diff(self, grad: Array) -> Array:
        """The bias-corrected Adam step, with the array ops of
        vulkpy/nn/optimizers.py:235-253 in their order"""
        o = self.opt
        # moments: m = b1 m + (1 - b1) g ; v = b2 v + (1 - b2) g^2
        self.m *= o.beta1
        self.m += (1 - o.beta1) * grad
        self.v *= o.beta2
        self.v += (1 - o.beta2) * (grad ** 2)
        self._advance_powers()
        # -lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)
        m_hat = self.m / (1 - self.beta1t)
        v_hat = self.v / (1 - self.beta2t)
        v_hat.sqrt(inplace=True)
        v_hat += o.eps
        m_hat *= (-o.lr)
        m_hat /= v_hat
        return m_hat

    def _update(self, value: Array, grad: Array):
        """grad2diff + ``value += diff`` as one kernel"""
        self._advance_powers()
        o = self.opt
        call("vkh_adam", value._gpu._handle, value._ptr, grad._ptr, self.m._ptr, self.v._ptr,
             value.size, float(o.lr), float(o.beta1), float(o.beta2), float(o.eps),
             float(self.beta1t), float(self.beta2t))
        value._set_job([grad])
        self.m._set_job()
        self.v._set_job()


class Adam(Optimizer):
    """
    Adam Optimizer (Kingma and Ba, ICLR 2015)

    Per element, with gradient g and step count t: exponential moving
    averages m <- beta1 m + (1 - beta1) g and v <- beta2 v + (1 - beta2) g^2,
    then the parameter moves by -lr * m' / (sqrt(v') + eps), where
    m' = m / (1 - beta1^t) and v' = v / (1 - beta2^t) undo the averages'
    bias toward their zero start.
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
