"""
Neural Network Layer Module (:mod:`vulkpy_amd.nn.layers`)
=========================================================

Same API as vulkpy/nn/layers.py.  Dense runs on the matrix cores: the forward
is one GEMM with the bias in its epilogue (instead of batch_affine.comp) and
the backward is two GEMMs plus a column sum (instead of materialising the
B x out x in product, layers.py:131-135).
"""
from __future__ import annotations
from typing import Callable, Iterable, Optional

from vulkpy_amd._lib import ACT, GEMM_MODE, call
from vulkpy_amd import vkarray
from vulkpy_amd.vkarray import GPU, Array
from .core import Module, Optimizer, Regularizer
from .parameters import Parameter
from .initializers import HeNormal


__all__ = ["Dense", "ReLU", "Sigmoid", "Softmax"]


class Dense(Module):
    """
    Fully connected Dense Layer
    """

    def __init__(self, gpu: GPU, input_dim: int, output_dim: int, *,
                 w_init: Optional[Callable[[GPU, Iterable[int]], Array]] = None,
                 b_init: Optional[Callable[[GPU, Iterable[int]], Array]] = None,
                 w_opt: Optional[Optimizer] = None,
                 b_opt: Optional[Optimizer] = None,
                 w_reg: Optional[Regularizer] = None,
                 b_reg: Optional[Regularizer] = None,
                 precision: Optional[str] = None):
        """
        Initialize Dense

        Parameters
        ----------
        gpu : vulkpy_amd.GPU
            GPU
        input_dim : int
            Input dimension
        output_dim : int
            Output dimension
        w_init, b_init, w_opt, b_opt, w_reg, b_reg
            As in vulkpy (HeNormal weights, zero bias, Adam by default).
        precision : {"fp32", "bf16", "bf16x3"}, optional
            Product precision of the three GEMMs; default is
            ``vulkpy_amd.get_matmul_precision()`` (``"fp32"``).
        """
        self.input_dim = int(input_dim)
        self.output_dim = int(output_dim)
        self.precision = precision

        if w_init is None:
            w_init = HeNormal(gpu, self.input_dim)

        self.w = Parameter(gpu, shape=(self.output_dim, self.input_dim),
                           initializer=w_init, opt=w_opt, regularizer=w_reg)
        self.b = Parameter(gpu, shape=(self.output_dim,),
                           initializer=b_init, opt=b_opt, regularizer=b_reg)

    def _mode(self) -> int:
        return GEMM_MODE[self.precision or vkarray.get_matmul_precision()]

    def forward(self, x: Array) -> Array:
        r"""
        Forward: :math:`y = x W^T + b` (batch_affine.comp:25-39)
        """
        B, K = int(x.shape[0]), int(x.shape[1])
        if K != self.input_dim:
            raise ValueError(f"Incompatible shapes: {x.shape} vs input_dim {self.input_dim}")
        y = Array(x._gpu, shape=(B, self.output_dim))
        call("vkh_gemm", x._gpu._handle, self._mode(), 0, 1, B, self.output_dim, K, 1.0,
             x._ptr, K, self.w.value._ptr, K, 0.0, y._ptr, self.output_dim,
             self.b.value._ptr, ACT["none"], None)
        y._set_job([self.w.value, self.b.value, x])
        return y

    def backward(self, dy: Array) -> Array:
        r"""
        Backward

        .. math::

            dx = dy W\\
            dW = dy ^T \cdot x\\
            db = \sum_b dy

        Gradients are accumulated into ``w.grad`` / ``b.grad``
        (``Parameter.add_grad``, vulkpy/nn/layers.py:126-136).
        """
        g = dy._gpu._handle
        B, O, I = int(dy.shape[0]), self.output_dim, self.input_dim
        x = self._x
        mode = self._mode()
        if self.b.grad is not None:
            call("vkh_bias_grad", g, dy._ptr, self.b.grad._ptr, B, O, 1.0)
            self.b.grad._set_job([dy])
        if self.w.grad is not None:
            # w.grad[O, I] += dy^T[O, B] x[B, I]
            call("vkh_gemm", g, mode, 1, 0, O, I, B, 1.0, dy._ptr, O, x._ptr, I,
                 1.0, self.w.grad._ptr, I, None, ACT["none"], None)
            self.w.grad._set_job([dy, x])
        # dx[B, I] = dy[B, O] W[O, I]
        dx = Array(dy._gpu, shape=(B, I))
        call("vkh_gemm", g, mode, 0, 0, B, I, O, 1.0, dy._ptr, O, self.w.value._ptr, I,
             0.0, dx._ptr, I, None, ACT["none"], None)
        dx._set_job([dy, self.w.value])
        return dx

    def zero_grad(self):
        """
        Clear accumulated gradients
        """
        self.w.zero_grad()
        self.b.zero_grad()

    def update(self):
        """
        Update values with accumulated gradients
        """
        self.w.update()
        self.b.update()


def _act_forward(x: Array, act: str) -> Array:
    y = Array(x._gpu, shape=x.shape)
    call("vkh_activation", x._gpu._handle, ACT[act], x._ptr, y._ptr, x.size)
    y._set_job([x])
    return y


def _act_backward(y: Array, dy: Array, act: str) -> Array:
    if tuple(y.shape) != tuple(dy.shape):
        raise ValueError(f"Incompatible shapes: {y.shape} vs {dy.shape}")
    dx = Array(dy._gpu, shape=dy.shape)
    call("vkh_activation_backward", dy._gpu._handle, ACT[act], y._ptr, dy._ptr, dx._ptr, dy.size)
    dx._set_job([y, dy])
    return dx


class ReLU(Module):
    """
    Rectified Linear Unit (ReLU)
    """
    def forward(self, x: Array) -> Array:
        r"""
        Forward: :math:`y = \max(x, 0)` (nn/layers.py:185)
        """
        return _act_forward(x, "relu")

    def backward(self, dy: Array) -> Array:
        r"""
        Backward: :math:`dx = dy \cdot \max(\rm{sign}(y), 0)` (nn/layers.py:207-210)
        """
        return _act_backward(self._y, dy, "relu")


class Sigmoid(Module):
    """
    Sigmoid
    """
    def forward(self, x: Array) -> Array:
        r"""
        Forward: :math:`y = 1/(1 + \exp (-x))` (nn/layers.py:240-244)
        """
        return _act_forward(x, "sigmoid")

    def backward(self, dy: Array) -> Array:
        r"""
        Backward: :math:`dx = dy \cdot y(1 - y)` (nn/layers.py:264-267)
        """
        return _act_backward(self._y, dy, "sigmoid")


class Softmax(Module):
    """
    SoftMax
    """
    def forward(self, x: Array) -> Array:
        r"""
        Forward: :math:`y = \exp (x - \max x) / \sum _i \exp(x_i - \max x)`
        over axis 1 (nn/layers.py:297-300)
        """
        if len(x.shape) == 2:
            y = Array(x._gpu, shape=x.shape)
            call("vkh_softmax", x._gpu._handle, x._ptr, y._ptr, int(x.shape[0]), int(x.shape[1]))
            y._set_job([x])
            return y
        X = x - x.maximum(axis=1, rebroadcast=True)
        X.exp(inplace=True)
        X /= X.sum(axis=1, rebroadcast=True)
        return X

    def backward(self, dy: Array) -> Array:
        r"""
        Backward: :math:`dx = dy \cdot y(1 - y)`, the Jacobian diagonal only,
        as in nn/layers.py:320-323
        """
        return _act_backward(self._y, dy, "sigmoid")
