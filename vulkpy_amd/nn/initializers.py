"""
Neural Network Initializer Module (:mod:`vulkpy_amd.nn.initializers`)
=====================================================================

Same as vulkpy/nn/initializers.py:16-89.
"""
from __future__ import annotations
from typing import Iterable, Optional

import numpy as np

from vulkpy_amd.vkarray import GPU, Array
from vulkpy_amd.random import Xoshiro128pp


__all__ = ["Constant", "HeNormal"]


class Initializer:
    def __call__(self, gpu: GPU, shape: Iterable[int]) -> Array:
        raise NotImplementedError


class Constant(Initializer):
    """
    Constant Initializer
    """
    def __init__(self, value: float):
        self.value = value

    def __call__(self, gpu: GPU, shape: Iterable[int]) -> Array:
        p = Array(gpu, shape=shape)
        p[:] = self.value
        return p


class HeNormal(Initializer):
    r"""
    He Normal Initializer

    Standard deviation is :math:`\sigma = \sqrt{2/d_{\text{in}}}`.
    """
    def __init__(self, gpu: GPU, input_dim: int, *, seed: Optional[int] = None):
        self.rng = Xoshiro128pp(gpu, seed=seed)
        self.stddev = np.sqrt(2 / input_dim)

    def __call__(self, gpu: GPU, shape: Iterable[int]):
        return self.rng.normal(shape=shape, stddev=self.stddev)
