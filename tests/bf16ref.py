"""NumPy statement of the library's FP32 -> BF16 rounding (the bits every BF16
operand and copy carries): round to nearest even, NaN kept NaN (quieted, top
bits kept), and a nonzero value that would round to zero keeps the smallest
BF16 magnitude of its sign."""
import numpy as np


def bf16_bits(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    mag = u & 0x7FFFFFFF
    b = np.where((b & 0x7FFF) == 0, b | (mag != 0), b)
    b = np.where(mag > 0x7F800000, (u >> 16) | 0x40, b)
    return (b & 0xFFFF).astype(np.uint16)
