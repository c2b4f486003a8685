"""
vulkpy: GPGPU arrays, random numbers and neural networks on one MI355X
======================================================================

The import name scripts written for vulkpy use (``import vulkpy as vk``,
``from vulkpy import nn``, ``from vulkpy.util import enable_debug``).  Every
module under this name IS the corresponding module of :mod:`vulkpy_amd`,
registered in ``sys.modules`` under the reference's import paths
(vulkpy/__init__.py:26-28 and its submodules), so both names share one set of
classes, one device context per GPU and one loaded ``libvulkpy_hip.so``:
``vulkpy.Array is vulkpy_amd.Array``.

Examples
--------
>>> import vulkpy as vk
>>> gpu = vk.GPU()
>>> a = vk.Array(gpu, data=[1, 2, 3])
>>> b = vk.Array(gpu, data=[3, 3, 3])
>>> c = a + b
>>> print(c)
[4. 5. 6.]
"""
import importlib
import sys

import vulkpy_amd as _impl

# Submodules published under the reference's paths.  Each is imported from
# the implementation package and the same module object is entered for the
# vulkpy name, so isinstance checks, module state and pickled class paths
# agree between the two names.
_SUBMODULES = (
    "vkarray", "vktyping", "random", "util",
    "nn", "nn.core", "nn.initializers", "nn.layers", "nn.losses",
    "nn.models", "nn.optimizers", "nn.parameters", "nn.regularizers", "nn.graph",
)

for _name in _SUBMODULES:
    _mod = importlib.import_module(f"{_impl.__name__}.{_name}")
    sys.modules[f"{__name__}.{_name}"] = _mod

from vulkpy_amd.vkarray import GPU, U32Array, Shape, Array, zeros  # noqa: E402
from vulkpy_amd import random, nn, util, vkarray, vktyping  # noqa: E402
from vulkpy_amd import (set_matmul_precision, get_matmul_precision,  # noqa: E402
                        gemm_configs, autotune_gemm, reduce_configs, autotune_reduce)

__all__ = ["GPU", "U32Array", "Shape", "Array", "zeros", "random", "nn", "util",
           "set_matmul_precision", "get_matmul_precision",
           "gemm_configs", "autotune_gemm", "reduce_configs", "autotune_reduce"]

del _name, _mod
