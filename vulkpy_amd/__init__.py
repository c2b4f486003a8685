"""
vulkpy on MI355X: GPGPU arrays on HIP (gfx950)
=============================================

Drop-in backend for the vulkpy API: the same ``GPU``, ``Array``,
``U32Array``, ``Shape``, ``zeros``, ``random`` and ``nn`` names and
signatures, executed by hand-written HIP kernels of ``libvulkpy_hip.so``.

Examples
--------
>>> import vulkpy_amd as vk
>>> gpu = vk.GPU()
>>> a = vk.Array(gpu, data=[1, 2, 3])
>>> b = vk.Array(gpu, data=[3, 3, 3])
>>> c = a + b
>>> print(c)
[4. 5. 6.]
"""
from .vkarray import (GPU, U32Array, Shape, Array, zeros,
                      set_matmul_precision, get_matmul_precision,
                      gemm_configs, autotune_gemm, reduce_configs, autotune_reduce)
from . import random
from . import nn
from . import util
