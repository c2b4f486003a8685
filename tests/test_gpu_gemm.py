"""GEMM on the matrix cores: layouts, modes and epilogues against FP64."""
import numpy as np
import pytest

import vulkpy_amd as vk
from vulkpy_amd import _lib

pytestmark = pytest.mark.gpu


def run_gemm(gpu, mode, ta, tb, A, B, C0=None, alpha=1.0, beta=0.0, bias=None, act="none"):
    """C = alpha op(A) op(B) + beta C0 + bias through vkh_gemm."""
    M = A.shape[1] if ta else A.shape[0]
    K = A.shape[0] if ta else A.shape[1]
    N = B.shape[0] if tb else B.shape[1]
    gA, gB = vk.Array(gpu, data=A), vk.Array(gpu, data=B)
    gC = vk.Array(gpu, data=C0) if C0 is not None else vk.Array(gpu, shape=(M, N))
    gb = vk.Array(gpu, data=bias) if bias is not None else None
    _lib.call("vkh_gemm", gpu._handle, _lib.GEMM_MODE[mode], int(ta), int(tb), M, N, K, alpha,
              gA._ptr, A.shape[1], gB._ptr, B.shape[1], beta, gC._ptr, N,
              gb._ptr if gb is not None else None, _lib.ACT[act], None)
    gC._set_job([gA, gB])
    return np.asarray(gC)


def ref_gemm(ta, tb, A, B):
    a = A.astype(np.float64).T if ta else A.astype(np.float64)
    b = B.astype(np.float64).T if tb else B.astype(np.float64)
    return a @ b, np.abs(a) @ np.abs(b)


TOL = {"fp32": 2.0 ** -20, "bf16": 2.0 ** -7, "bf16x3": 2.0 ** -15}


@pytest.mark.parametrize("mode", ["fp32", "bf16", "bf16x3"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(128, 128, 64), (130, 70, 33), (1, 257, 300), (300, 1, 17), (5, 6, 1)])
def test_layouts_modes(gpu, mode, ta, tb, M, N, K):
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    got = run_gemm(gpu, mode, ta, tb, A, B)
    ref, bound = ref_gemm(ta, tb, A, B)
    err = (np.abs(got - ref) / bound).max()
    assert err <= TOL[mode], err


def test_output_layout_exact(gpu):
    # integer-valued operands give exact products and sums in every mode,
    # so any misplaced accumulator element shows up as a mismatch
    rng = np.random.default_rng(0)
    A = rng.integers(-4, 5, (256, 96)).astype(np.float32)
    B = rng.integers(-4, 5, (96, 192)).astype(np.float32)
    want = A @ B
    for mode in ("fp32", "bf16", "bf16x3"):
        got = run_gemm(gpu, mode, 0, 0, A, B)
        np.testing.assert_array_equal(got, want)


def test_epilogue(gpu):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((70, 40)).astype(np.float32)
    B = rng.standard_normal((50, 40)).astype(np.float32)
    C0 = rng.standard_normal((70, 50)).astype(np.float32)
    bias = rng.standard_normal(50).astype(np.float32)
    got = run_gemm(gpu, "fp32", 0, 1, A, B, C0=C0, alpha=0.5, beta=2.0, bias=bias)
    ref = 0.5 * (A.astype(np.float64) @ B.T.astype(np.float64)) + 2.0 * C0 + bias
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-4)
    got = run_gemm(gpu, "fp32", 0, 1, A, B, bias=bias, act="relu")
    ref = np.maximum(A.astype(np.float64) @ B.T.astype(np.float64) + bias, 0)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-4)


def test_split_bf16_accuracy(gpu):
    rng = np.random.default_rng(2)
    A = rng.standard_normal((512, 1024)).astype(np.float32)
    B = rng.standard_normal((1024, 512)).astype(np.float32)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    errs = {}
    for mode in ("fp32", "bf16", "bf16x3"):
        got = run_gemm(gpu, mode, 0, 0, A, B)
        errs[mode] = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert errs["bf16"] <= 1e-2
    assert errs["bf16x3"] <= 1e-5
    assert errs["fp32"] <= 1e-6
