"""
Random Module (:mod:`vulkpy_amd.random`)
========================================

Device-side pseudo random numbers with the interface of vulkpy/random.py:
one generator object holds ``size`` xoshiro128++ states in device memory, and
every draw is a single kernel over them (the state of each stream stays in
registers for all of its outputs).

Examples
--------
>>> import vulkpy_amd as vk
>>> gpu = vk.GPU()
>>> r = vk.random.Xoshiro128pp(gpu, seed=0)

Uniform floats in [0, 1) (the reference's own example values for seed 0):

>>> print(r.random(shape=(3,)))
[0.42977667 0.8235899  0.90622926]

Standard normal floats, continuing the same streams:

>>> print(r.normal(shape=(3,)))
[-2.3403292  0.7247794  0.7118352]
"""

from __future__ import annotations

import ctypes
import functools
import os
from typing import cast, Iterable, Optional

import numpy as np

from . import _lib
from ._lib import call
from . import vkarray as vk

__all__ = ["Xoshiro128pp"]


class PRNG(vk.Resource):
    """Generator interface (vulkpy/random.py:40-189).  A subclass supplies
    ``random`` and ``randint``; ``normal`` and ``randrange`` here are built on
    ``random`` (the Box-Muller and scaling kernels)."""

    _2p32 = int(2 ** 32)

    def __init__(self, gpu: vk.GPU):
        self._gpu = gpu

    def random(self, *,
               shape: Optional[Iterable[int]] = None,
               buffer: Optional[vk.Array] = None) -> vk.Array:
        raise NotImplementedError

    def randint(self, *,
                shape: Optional[Iterable[int]] = None,
                buffer: Optional[vk.U32Array] = None) -> vk.U32Array:
        raise NotImplementedError

    @staticmethod
    def _out(gpu, shape, buffer, cls):
        if buffer is None:
            if shape is None:
                raise ValueError("One of `shape` and  `buffer` must be specified.")
            return cls(gpu, shape=shape)
        # earlier work may still write the output buffer
        buffer.wait()
        return buffer

    def normal(self, *,
               shape: Optional[Iterable[int]] = None,
               buffer: Optional[vk.Array] = None,
               mean: float = 0.0,
               stddev: float = 1.0) -> vk.Array:
        """
        Normal draws with ``mean`` and ``stddev``

        Parameters
        ----------
        shape : iterable of ints, optional
            Shape of a new output array.
        buffer : vulkpy_amd.Array, optional
            Existing output array (used instead of ``shape``).

        Returns
        -------
        vulkpy_amd.Array
            The output, filled when its job completes.

        Raises
        ------
        ValueError
            If neither ``shape`` nor ``buffer`` is given

        Notes
        -----
        Draws n uniforms (n even) or n + 1 (n odd) and transforms pairs with
        the Box-Muller method (vulkpy/random.py:60-124): element 2i is
        ``mean + r sin(theta)``, element 2i+1 is ``mean + r cos(theta)``.
        """
        buffer = cast(vk.Array, self._out(self._gpu, shape, buffer, vk.Array))
        n = buffer.size
        nu = n + (n % 2)
        u = vk.Array(self._gpu, shape=(nu,))
        u = self.random(buffer=u)
        call("vkh_box_muller", self._gpu._handle, u._ptr, buffer._ptr, n,
             float(mean), float(stddev))
        buffer._set_job([u])
        return buffer

    def randrange(self, *,
                  shape: Optional[Iterable[int]] = None,
                  buffer: Optional[vk.U32Array] = None,
                  low: int = 0,
                  high: int = int(2 ** 32)) -> vk.U32Array:
        """
        Integers drawn from [low, high)

        Raises
        ------
        ValueError
            If neither ``shape`` nor ``buffer`` is given, or unless
            0 <= low < high <= 2^32.
        """
        self._check_range(low, high)
        if (low == 0) and (high == self._2p32):
            return self.randint(shape=shape, buffer=buffer)
        raise NotImplementedError

    def _check_range(self, low, high):
        if low < 0:
            raise ValueError(f"`low` must be non negative integer, but {low}")
        if high > self._2p32:
            raise ValueError(f"`high` must not be greater than 2^32, but {high}")
        if low >= high:
            raise ValueError(f"`low` must be smaller than `high`, but {low}, {high}")

    def wait(self):
        pass


class Xoshiro128pp(PRNG):
    """
    ``size`` parallel xoshiro128++ streams (Blackman and Vigna's 32-bit
    generator with the ++ scrambler) in device memory

    The initial states are those vulkpy computes (_vkarray.cc:643-675): a
    splitmix64 chain from the seed gives stream 0, and stream i is stream
    i - 1 advanced by the reference's jump as written.  Here they are formed
    in parallel from GF(2) matrix powers of that jump, with identical bits.
    A draw of n values assigns output k to stream ``k % size``, step
    ``k // size`` (_vkarray.cc:697-717), in one kernel.
    """

    def __init__(self, gpu: vk.GPU, size: int = 64, *, seed: Optional[int] = None):
        """
        Parameters
        ----------
        gpu : vulkpy_amd.GPU
            Device holding the streams' states.
        size : int
            Number of streams (default 64).
        seed : int, optional
            Seed of the splitmix64 chain; ``None`` (default) takes 32 bits
            from the operating system's entropy source.
        """
        super().__init__(gpu)
        if seed is None:
            # std::random_device yields 32 bits (_vkarray.cc:676-679)
            seed = int.from_bytes(os.urandom(4), "little")
        if not isinstance(seed, (int, np.integer)) or not (0 <= int(seed) < 2 ** 64):
            raise TypeError(f"seed must be an integer in [0, 2^64), got {seed!r}")
        if not (1 <= int(size) < 2 ** 32):
            raise ValueError(f"size must be in [1, 2^32), got {size}")
        out = ctypes.c_void_p()
        call("vkh_xoshiro_create", gpu._handle, int(size), int(seed), ctypes.byref(out))
        self._rng = out.value
        self.size = int(size)

    def __del__(self):
        try:
            if self._rng:
                rng, self._rng = self._rng, None
                # the release waits for the stream: not during a graph capture
                if not vk._after_capture(functools.partial(_lib.lib.vkh_xoshiro_destroy, rng)):
                    _lib.lib.vkh_xoshiro_destroy(rng)
        except Exception:
            pass

    def random(self, *,
               shape: Optional[Iterable[int]] = None,
               buffer: Optional[vk.Array] = None) -> vk.Array:
        """
        Uniform floats in [0, 1): the top 23 bits of each draw as a mantissa

        Raises
        ------
        ValueError
            If neither ``shape`` nor ``buffer`` is given.
        """
        buffer = cast(vk.Array, self._out(self._gpu, shape, buffer, vk.Array))
        call("vkh_xoshiro_float", self._rng, buffer._ptr, buffer.size)
        buffer._set_job([self])
        return buffer

    def randint(self, *,
                shape: Optional[Iterable[int]] = None,
                buffer: Optional[vk.U32Array] = None) -> vk.U32Array:
        """
        Raw 32-bit draws (uint32 in [0, 2^32))

        Raises
        ------
        ValueError
            If neither ``shape`` nor ``buffer`` is given.
        """
        buffer = cast(vk.U32Array, self._out(self._gpu, shape, buffer, vk.U32Array))
        call("vkh_xoshiro_uint32", self._rng, buffer._ptr, buffer.size)
        buffer._set_job([self])
        return buffer

    def normal(self, *,
               shape: Optional[Iterable[int]] = None,
               buffer: Optional[vk.Array] = None,
               mean: float = 0.0,
               stddev: float = 1.0) -> vk.Array:
        """
        Normal draws: the uniforms and their Box-Muller transform in one
        kernel, with the values and state advance of ``PRNG.normal``.
        """
        buffer = cast(vk.Array, self._out(self._gpu, shape, buffer, vk.Array))
        call("vkh_xoshiro_normal", self._rng, buffer._ptr, buffer.size,
             float(mean), float(stddev))
        buffer._set_job([self])
        return buffer

    def randrange(self, *,
                  shape: Optional[Iterable[int]] = None,
                  buffer: Optional[vk.U32Array] = None,
                  low: int = 0,
                  high: int = int(2 ** 32)) -> vk.U32Array:
        """
        Integers from [low, high) as ``low + uint(range * u)`` of a uniform
        drawn in the same kernel (prng_randrange.comp:24-26)

        Raises
        ------
        ValueError
            If neither ``shape`` nor ``buffer`` is given, or unless
            0 <= low < high <= 2^32.
        """
        self._check_range(low, high)
        if (low == 0) and (high == self._2p32):
            return self.randint(shape=shape, buffer=buffer)
        buffer = cast(vk.U32Array, self._out(self._gpu, shape, buffer, vk.U32Array))
        call("vkh_xoshiro_randrange", self._rng, buffer._ptr, buffer.size,
             int(low), int(high - 1))
        buffer._set_job([self])
        return buffer

    def state(self) -> np.ndarray:
        """Copy of the [size, 4] uint32 state (waits for pending work)."""
        ptr = ctypes.c_void_p()
        call("vkh_xoshiro_state", self._rng, ctypes.byref(ptr))
        out = np.empty((self.size, 4), dtype=np.uint32)
        call("vkh_download", self._gpu._handle, out.ctypes.data, ptr.value, out.nbytes)
        return out
