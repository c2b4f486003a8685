"""
Neural Network Model Module (:mod:`vulkpy_amd.nn.models`)
=========================================================

``Sequence`` with the interface of vulkpy/nn/models.py:15-105.  One training
step runs these stages in this order (the order the reference's ``train``
fixes, models.py:71-76):

1. forward through the layers and the loss;
2. zero every layer's gradients;
3. backward: the loss gradient through the layers in reverse;
4. (``regularize=True`` only) regularizer gradients of every parameter;
5. every layer's update.

Each stage only enqueues kernels on the device's stream, so a step issues no
host wait.  ``nn.GraphSequence`` runs the same stages, for Dense / ReLU /
Sigmoid networks, as one captured hipGraph.
"""
from __future__ import annotations
from typing import Iterable, Iterator, Optional, Tuple, Union

from vulkpy_amd.vkarray import Array
from .core import Module, Loss
from .parameters import Parameter


__all__ = ["Sequence"]


class Sequence:
    """
    Sequential Model: layers applied one after the other, then a loss

    Parameters
    ----------
    layers : iterable of vulkpy_amd.nn.Module
        Layers in the order they are applied
    loss : vulkpy_amd.nn.Loss
        Loss of the last layer's output against the targets
    regularize : bool, optional
        If ``True``, ``train`` adds every parameter's regularizer gradient
        (``Parameter.regular_grad``) between backward and update.  The
        default ``False`` is the reference's behaviour, where regularizers
        never enter training (vulkpy/nn/models.py:55-78).
    """
    def __init__(self,
                 layers: Iterable[Module],
                 loss: Loss, *,
                 regularize: bool = False):
        self.L: Tuple[Module, ...] = tuple(layers)
        self.loss: Loss = loss
        self.regularize = regularize

    def _apply_layers(self, x: Array) -> Array:
        h = x
        for layer in self.L:
            h = layer(h)
        return h

    def _parameters(self) -> Iterator[Parameter]:
        for layer in self.L:
            for attr in vars(layer).values():
                if isinstance(attr, Parameter):
                    yield attr

    def _learn(self):
        """stages 2-5 for the loss of the last forward"""
        for layer in self.L:
            layer.zero_grad()
        g = self.loss.grad()
        for layer in reversed(self.L):
            g = layer.backward(g)
        if self.regularize:
            for p in self._parameters():
                p.regular_grad()
        for layer in self.L:
            layer.update()

    def train(self, x: Array, y: Array) -> Tuple[Array, Array]:
        """
        One training step on the batch ``(x, y)``

        Returns
        -------
        y : vulkpy_amd.Array
            Prediction of the batch (before this step's update)
        loss : vulkpy_amd.Array
            Loss of that prediction
        """
        pred = self._apply_layers(x)
        loss = self.loss(pred, y)
        self._learn()
        return pred, loss

    def predict(self,
                x: Array,
                y: Optional[Array] = None) -> Union[Array, Tuple[Array, Array]]:
        """
        Prediction of ``x``, and its loss against ``y`` when ``y`` is given
        """
        pred = self._apply_layers(x)
        if y is None:
            return pred
        return pred, self.loss(pred, y)
