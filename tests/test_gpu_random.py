"""GPU parity of Xoshiro128++ (F18): bit-exact uint32 / float streams and state
against the oracle (which reproduces the reference's KAT, random.py:17-24),
Box-Muller and randrange against the oracle, and the reference's own tests
(test/test_random.py)."""
import numpy as np
import pytest

import vulkpy_amd as vk
from vulkpy_amd._lib import call
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_docstring_kat(gpu):
    r = vk.random.Xoshiro128pp(gpu, seed=0)
    np.testing.assert_array_equal(np.asarray(r.random(shape=(3,))),
                                  np.array([0.42977667, 0.8235899, 0.90622926], dtype=np.float32))
    np.testing.assert_allclose(r.normal(shape=(3,)), [-2.3403292, 0.7247794, 0.7118352], rtol=1e-6)


@pytest.mark.parametrize("size", [1, 2, 3, 64, 100, 4096, 5000])
def test_seeding_matches_oracle(gpu, size):
    r = vk.random.Xoshiro128pp(gpu, size=size, seed=12345)
    np.testing.assert_array_equal(r.state(), O.xoshiro_seed(size, 12345))


@pytest.mark.parametrize("size,calls", [
    (64, [3, 10, 64, 65, 200, 1]),
    (7, [1, 6, 7, 8, 50]),
    (1000, [999, 1000, 1001, 4321, 17]),
    # multiples of 256 states with a multiple of 4 steps per state take the
    # LDS-staged row kernel; the 1000 in between leaves the states uneven
    (512, [2048, 4096, 6144, 1000, 2048, 512 * 8]),
    (4096, [4096 * 64, 4096 * 8]),
])
def test_stream_interleaved(gpu, size, calls):
    r = vk.random.Xoshiro128pp(gpu, size=size, seed=7)
    o = O.Xoshiro128pp(size, 7)
    for i, n in enumerate(calls):
        if i % 2 == 0:
            np.testing.assert_array_equal(np.asarray(r.randint(shape=(n,))), o.randint(n))
        else:
            got = np.asarray(r.random(shape=(n,)))
            np.testing.assert_array_equal(got.view(np.uint32), o.random(n).view(np.uint32))
    np.testing.assert_array_equal(r.state(), o.state)


@pytest.mark.parametrize("size", [64, 63, 2, 1])
@pytest.mark.parametrize("n", [1, 2, 10, 11, 64, 65, 333])
def test_normal_matches_oracle(gpu, size, n):
    r = vk.random.Xoshiro128pp(gpu, size=size, seed=3)
    o = O.Xoshiro128pp(size, 3)
    r.random(shape=(5,)); o.random(5)
    got = np.asarray(r.normal(shape=(n,), mean=0.5, stddev=2.0))
    want = o.normal(n, 0.5, 2.0)
    # The kernel evaluates log / sin / cos in FP32 device math (within 2 ulp);
    # the oracle uses correctly rounded values.  The last step, mean + r*s,
    # can cancel, so the bound is 4 ulp of its larger operand.
    scale = np.abs(want - np.float32(0.5)) + np.float32(0.5)
    assert np.all(np.abs(got - want) <= 4 * np.spacing(scale)), np.abs(got - want).max()
    np.testing.assert_array_equal(r.state(), o.state)


def test_fused_normal_equals_two_pass(gpu):
    """The fused kernel evaluates ln(1 - u) and sqrt on their generated
    ranges (hardware log2 / sqrt with the device functions' own steps and no
    range handling); the two-pass path (uniforms, then vkh_box_muller with the
    device logf / sqrtf) must give the same bits.  2^25 pairs cover nearly
    all 2^23 float uniforms u1 can take."""
    n = 1 << 26
    fused = np.asarray(vk.random.Xoshiro128pp(gpu, size=1 << 16, seed=7).normal(shape=(n,), mean=0.25, stddev=1.5))
    u = vk.random.Xoshiro128pp(gpu, size=1 << 16, seed=7).random(shape=(n,))
    two = vk.Array(gpu, shape=(n,))
    call("vkh_box_muller", gpu._handle, u._ptr, two._ptr, n, 0.25, 1.5)
    two._set_job([u])
    np.testing.assert_array_equal(fused.view(np.uint32), np.asarray(two).view(np.uint32))
    # odd length: the last pair's second value is dropped, the rest agree
    odd = np.asarray(vk.random.Xoshiro128pp(gpu, size=1 << 16, seed=7).normal(shape=(n - 1,), mean=0.25, stddev=1.5))
    np.testing.assert_array_equal(odd.view(np.uint32), fused[:-1].view(np.uint32))


def test_randrange(gpu):
    r = vk.random.Xoshiro128pp(gpu, seed=4)
    np.testing.assert_array_equal(np.asarray(r.randrange(shape=(5,), low=3, high=4)), [3] * 5)
    r = vk.random.Xoshiro128pp(gpu, size=64, seed=9)
    o = O.Xoshiro128pp(64, 9)
    for low, high in ((0, 10), (5, 1000), (0, 2 ** 32), (7, 2 ** 31)):
        got = np.asarray(r.randrange(shape=(300,), low=low, high=high))
        np.testing.assert_array_equal(got, o.randrange(300, low, high))
    with pytest.raises(ValueError):
        r.randrange(shape=(3,), low=-1, high=3)
    with pytest.raises(ValueError):
        r.randrange(shape=(3,), low=5, high=5)
    with pytest.raises(ValueError):
        r.randrange(shape=(3,), low=0, high=2 ** 32 + 1)


def test_reference_random_tests(gpu):
    rng = vk.random.Xoshiro128pp(gpu)
    for shape in ((3,), (17,), (65,), (5, 5, 5)):
        a = np.asarray(rng.random(shape=shape))
        assert a.shape == shape and (a >= 0).all() and (a < 1).all()
    b = vk.Array(gpu, shape=(5,))
    b = rng.random(buffer=b)
    assert (np.asarray(b) < 1).all()
    a = rng.randint(shape=(5,))
    assert tuple(a.shape) == (5,)
    for n in (10, 11):
        r1 = vk.random.Xoshiro128pp(gpu, seed=0)
        r2 = vk.random.Xoshiro128pp(gpu, seed=0)
        a1 = r1.normal(shape=(n,))
        a2 = r2.normal(shape=(n,), mean=5, stddev=3)
        np.testing.assert_allclose((a2 - 5) / a1, np.full((n,), 3), rtol=1e-5)
    with pytest.raises(ValueError):
        rng.random()
    s1 = np.asarray(vk.random.Xoshiro128pp(gpu, seed=0).random(shape=(5,)))
    s2 = np.asarray(vk.random.Xoshiro128pp(gpu, seed=0).random(shape=(5,)))
    np.testing.assert_array_equal(s1, s2)


def test_large_state_seeding_parallel(gpu):
    # 2^16 states exercise the multi-threaded GF(2) seeding; compare to the
    # oracle's sequential jumps
    r = vk.random.Xoshiro128pp(gpu, size=1 << 16, seed=1)
    np.testing.assert_array_equal(r.state(), O.xoshiro_seed(1 << 16, 1))


@pytest.mark.parametrize("size", [1, 2, 3, 64, 1000])
def test_few_states_chunked_streams(gpu, size):
    """Few states and long streams take the chunked kernels (each state's
    chain split into chunks started by GF(2) jump-ahead): outputs, the state
    afterwards and the continuation must equal the sequential generator."""
    r = vk.random.Xoshiro128pp(gpu, size=size, seed=17)
    o = O.Xoshiro128pp(size, 17)
    n = (1 << 20) + 12345
    np.testing.assert_array_equal(np.asarray(r.randint(shape=(n,))), o.randint(n))
    np.testing.assert_array_equal(r.state(), o.state)
    got = np.asarray(r.random(shape=(n - 777,)))
    np.testing.assert_array_equal(got.view(np.uint32), o.random(n - 777).view(np.uint32))
    np.testing.assert_array_equal(r.state(), o.state)
    got = np.asarray(r.randrange(shape=(n // 2,), low=5, high=1000))
    np.testing.assert_array_equal(got, o.randrange(n // 2, 5, 1000))
    for m in (n, n + 1):  # even and odd counts (odd: n + 1 uniforms)
        got = np.asarray(r.normal(shape=(m,), mean=0.5, stddev=2.0))
        want = o.normal(m, 0.5, 2.0)
        scale = np.abs(want - np.float32(0.5)) + np.float32(0.5)
        assert np.all(np.abs(got - want) <= 4 * np.spacing(scale)), np.abs(got - want).max()
        np.testing.assert_array_equal(r.state(), o.state)
    # short calls after the long ones continue the same streams
    np.testing.assert_array_equal(np.asarray(r.randint(shape=(33,))), o.randint(33))
    np.testing.assert_array_equal(r.state(), o.state)


def test_henormal_shape_chunked(gpu):
    """The reference initializer's configuration (64 states, one call of
    in x out normals) at the C5 layer size."""
    r = vk.random.Xoshiro128pp(gpu, size=64, seed=2)
    o = O.Xoshiro128pp(64, 2)
    n = 4096 * 4096
    got = np.asarray(r.normal(shape=(n,), stddev=0.02))
    want = o.normal(n, 0.0, 0.02)
    assert np.all(np.abs(got - want) <= 4 * np.spacing(np.abs(want) + np.float32(0.0))
                  + 4 * np.spacing(np.float32(0.02))), np.abs(got - want).max()
    np.testing.assert_array_equal(r.state(), o.state)
