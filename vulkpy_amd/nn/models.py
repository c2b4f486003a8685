"""
Neural Network Model Module (:mod:`vulkpy_amd.nn.models`)
=========================================================

Same as vulkpy/nn/models.py:15-105.
"""
from __future__ import annotations
from typing import Iterable, Optional, Tuple, Union

from vulkpy_amd.vkarray import Array
from .core import Module, Loss
from .parameters import Parameter


__all__ = ["Sequence"]


class Sequence:
    """
    Sequential Model

    All layers sequentially connceted.
    """
    def __init__(self,
                 layers: Iterable[Module],
                 loss: Loss, *,
                 regularize: bool = False):
        """
        Parameters
        ----------
        layers : iterable of vulkpy_amd.nn.Module
            Layers to be called sequentially
        loss : vulkpy_amd.nn.Loss
            Loss layer
        regularize : bool, optional
            If ``True``, ``train`` adds every parameter's regularizer
            gradient (``Parameter.regular_grad``) between backward and update.
            The default ``False`` is the reference's behaviour, where
            regularizers never enter training (vulkpy/nn/models.py:55-78).
        """
        self.L: Tuple[Module, ...] = tuple(layers)
        self.loss: Loss = loss
        self.regularize = regularize

    def _forward(self, x: Array) -> Array:
        for _L in self.L:
            x = _L(x)
        return x

    def _backward(self):
        dx = self.loss.grad()
        for _L in self.L[::-1]:
            dx = _L.backward(dx)

    def _zero_grad(self):
        for _L in self.L:
            _L.zero_grad()

    def _update(self):
        for _L in self.L:
            _L.update()

    def _regularize(self):
        for _L in self.L:
            for p in vars(_L).values():
                if isinstance(p, Parameter):
                    p.regular_grad()

    def train(self, x: Array, y: Array) -> Tuple[Array, Array]:
        """
        Train model: forward, loss, zero_grad, backward, update
        (vulkpy/nn/models.py:55-78)

        Returns
        -------
        y : vulkpy_amd.Array
            Predicted Labels/Targets
        loss : vulkpy_amd.Array
            Loss
        """
        _y = self._forward(x)
        _loss = self.loss(_y, y)

        self._zero_grad()
        self._backward()
        if self.regularize:
            self._regularize()
        self._update()

        return _y, _loss

    def predict(self,
                x: Array,
                y: Optional[Array] = None) -> Union[Array, Tuple[Array, Array]]:
        """
        Predict Label/Target (and the loss when ``y`` is given)
        """
        _y = self._forward(x)
        if y is None:
            return _y

        _loss = self.loss(_y, y)
        return _y, _loss
