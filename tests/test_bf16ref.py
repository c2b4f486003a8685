"""The NumPy statement of the library's BF16 rounding against hand-worked
values (no GPU)."""
import numpy as np

from bf16ref import bf16_bits


def test_bf16_rounding_known_values():
    x = np.array([0.0, -0.0, 1.0, -1.5, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20,
                  2 ** -133, 2 ** -134, 1e-45, -1e-45, 2 ** -126, np.inf, -np.inf, 3.4e38],
                 dtype=np.float32)
    want = [0x0000, 0x8000, 0x3F80, 0xBFC0,
            0x3F80,            # a tie rounds to the even neighbour
            0x3F82,            # a tie between 0x3F81 and 0x3F82 rounds up to even
            0x3F81,            # above the tie: up
            0x0001,            # the smallest BF16 denormal
            0x0001, 0x0001, 0x8001,   # nonzero values that would round to zero
            0x0080, 0x7F80, 0xFF80,
            0x7F80]            # beyond the largest BF16 after rounding: inf
    assert [int(v) for v in bf16_bits(x)] == want
    nan = np.array([0x7FC00001, 0xFF800001], dtype=np.uint32).view(np.float32)
    b = bf16_bits(nan)
    assert all((int(v) & 0x7F80) == 0x7F80 and (int(v) & 0x7F) != 0 for v in b)  # still NaN
