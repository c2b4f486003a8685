"""GEMM on the matrix cores: layouts, modes and epilogues against FP64."""
import numpy as np
import pytest
from bf16ref import bf16_bits

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


# Skinny products (min(M, N) <= 16) take the FP32 vector path in every mode:
# row-dot when the long operand is K-contiguous (vectorised when K % 4 == 0
# and rows are 16-byte aligned), column walk with k chunks otherwise.
SKINNY = [(4096, 10, 1024), (16, 300, 777), (2000, 16, 4096), (1, 4096, 16384),
          (16384, 1, 4096), (3, 500, 5000), (700, 7, 3), (9, 9, 64)]


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", SKINNY)
def test_skinny_paths(gpu, ta, tb, M, N, K):
    rng = np.random.default_rng(M + 3 * N + 7 * K + 11 * ta + 13 * tb)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    ref, bound = ref_gemm(ta, tb, A, B)
    # classical bound of a length-K FP32 dot product: K * 2^-24 * (|A||B|)
    tol = max(2.0 ** -20, K * 2.0 ** -24)
    for mode in ("fp32", "bf16"):
        got = run_gemm(gpu, mode, ta, tb, A, B)
        err = (np.abs(got - ref) / bound).max()
        assert err <= tol, (mode, err)


def rowdot_order(X, w):
    """The row-dot's summation order in FP32: lane l adds x[k] * w[k] for
    k = 4l + 256i + j (i, then j = 0..3) from 0, then the wave's butterfly
    (v_l += v_(l xor o) for o = 32, 16, ..., 1); lane 0's value."""
    P, K = X.shape
    lanes = np.zeros((P, 64), dtype=np.float32)
    for i in range(K // 256):
        for j in range(4):
            k = np.arange(64) * 4 + 256 * i + j
            lanes = lanes + (X[:, k] * w[k]).astype(np.float32)
    o = 32
    while o >= 1:
        lanes = lanes + lanes[:, np.arange(64) ^ o]
        o //= 2
    return lanes[:, 0]


@pytest.mark.parametrize("K", [4096, 1280, 1024, 512])
def test_rowdot_summation_order(gpu, K):
    """The one-output row-dot (the MLP's last forward layer) sums in its
    documented order, whichever kernel takes the shape (weights held in
    registers for K = 1024..4096, streamed otherwise): bit for bit against a
    NumPy restatement of that order."""
    P = 67
    rng = np.random.default_rng(K)
    X = rng.standard_normal((P, K)).astype(np.float32)
    w = rng.standard_normal((1, K)).astype(np.float32)
    got = run_gemm(gpu, "fp32", 0, 1, X, w)[:, 0]
    np.testing.assert_array_equal(got.view(np.uint32), rowdot_order(X, w[0]).view(np.uint32))


@pytest.mark.parametrize("M,N,K,ta,tb", [(300, 10, 64, 0, 1), (10, 300, 64, 1, 0), (257, 1, 33, 1, 1)])
def test_skinny_epilogue(gpu, M, N, K, ta, tb):
    rng = np.random.default_rng(4)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    ref, _ = ref_gemm(ta, tb, A, B)
    got = run_gemm(gpu, "bf16", ta, tb, A, B, C0=C0, alpha=0.5, beta=2.0, bias=bias)
    np.testing.assert_allclose(got, 0.5 * ref + 2.0 * C0 + bias, rtol=1e-5, atol=1e-4)
    got = run_gemm(gpu, "fp32", ta, tb, A, B, bias=bias, act="sigmoid")
    np.testing.assert_allclose(got, 1 / (1 + np.exp(-(ref + bias))), rtol=1e-5, atol=1e-6)
    # aux: the activation-derivative epilogue of vkh_gemm_ex (ReLU)
    aux = np.maximum(rng.standard_normal((M, N)), 0).astype(np.float32)
    gA, gB, gX = vk.Array(gpu, data=A), vk.Array(gpu, data=B), vk.Array(gpu, data=aux)
    gC = vk.Array(gpu, shape=(M, N))
    _lib.call("vkh_gemm_ex", gpu._handle, _lib.GEMM_MODE["fp32"], ta, tb, M, N, K, 1.0,
              gA._ptr, A.shape[1], gB._ptr, B.shape[1], 0.0, gC._ptr, N, None, 0, None,
              gX._ptr, N, _lib.ACT["relu"])
    gC._set_job([gA, gB, gX])
    np.testing.assert_allclose(np.asarray(gC), ref * (aux > 0), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("mode,ta,tb,M,N,K", [
    (mode, ta, tb, M, N, K)
    for mode in ("bf16", "bf16x3") for ta, tb in ((0, 1), (1, 0))
    for M, N, K in ((300, 260, 200), (300, 200, 8192))
] + [
    # FP32: every layout, 16-byte runs (first shape), element loads (odd
    # extents, second shape) and k split into 8 chunks (third shape)
    ("fp32", ta, tb, M, N, K)
    for ta in (0, 1) for tb in (0, 1)
    for M, N, K in ((300, 260, 200), (301, 259, 203), (300, 200, 8192))
])
def test_configurations_bit_identical(gpu, mode, ta, tb, M, N, K):
    """Every kernel configuration accumulates each output over k in the
    same order, so tuning never changes a result (in BF16 modes the shape
    with K = 8192 splits k into 8 chunks, whose bounds do not depend on the
    configuration)."""
    rng = np.random.default_rng(9)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    cfgs = vk.gemm_configs(mode)
    assert len(cfgs) >= 2
    outs = []
    for i in range(len(cfgs)):
        _lib.call("vkh_gemm_set_config", gpu._handle, _lib.GEMM_MODE[mode], ta, tb, M, N, K, i)
        outs.append(run_gemm(gpu, mode, ta, tb, A, B))
    _lib.call("vkh_gemm_set_config", gpu._handle, _lib.GEMM_MODE[mode], ta, tb, M, N, K, -1)
    for i, o in enumerate(outs[1:], 1):
        np.testing.assert_array_equal(o, outs[0], err_msg=f"configuration {i} {cfgs[i]}")
    ref, bound = ref_gemm(ta, tb, A, B)
    assert (np.abs(outs[0] - ref) / bound).max() <= TOL[mode]


def test_autotune_api(gpu):
    import ctypes
    r = vk.autotune_gemm(gpu, 512, 384, 256, precision="bf16", trans_b=True, reps=2)
    assert 0 <= r["best"] < len(vk.gemm_configs("bf16"))
    assert all(t > 0 for t in r["ms"]) and r["ms"][r["best"]] == min(r["ms"])
    idx, tuned = ctypes.c_int(), ctypes.c_int()
    _lib.call("vkh_gemm_get_config", gpu._handle, _lib.GEMM_MODE["bf16"], 0, 1, 512, 384, 256,
              ctypes.byref(idx), ctypes.byref(tuned))
    assert (idx.value, tuned.value) == (r["best"], 1)
    r = vk.autotune_gemm(gpu, 640, 512, 384, precision="fp32", trans_a=True, reps=2)
    assert len(r["ms"]) == len(vk.gemm_configs("fp32")) >= 2
    assert all(t > 0 for t in r["ms"]) and r["ms"][r["best"]] == min(r["ms"])
    _lib.call("vkh_gemm_set_config", gpu._handle, _lib.GEMM_MODE["fp32"], 1, 0, 640, 512, 384, -1)
    # skinny shapes have a single kernel choice
    assert vk.autotune_gemm(gpu, 4096, 1, 512, precision="bf16")["ms"][0] == -1
    with pytest.raises(ValueError):
        _lib.call("vkh_gemm_set_config", gpu._handle, _lib.GEMM_MODE["bf16"], 0, 0, 8, 8, 8, 99)


@pytest.mark.parametrize("mode", ["bf16", "fp32"])
def test_autotune_ext_times_the_call(gpu, mode):
    """vkh_gemm_autotune_ext runs the given product (bias, ReLU, BF16 copies)
    in every configuration, keeps the fastest, and the product then gives the
    bits it gave before tuning."""
    import ctypes
    h = gpu._handle
    md = _lib.GEMM_MODE[mode]
    M, N, K = 640, 512, 384
    rng = np.random.default_rng(53)
    A = vk.Array(gpu, data=rng.standard_normal((M, K)).astype(np.float32))
    B = vk.Array(gpu, data=rng.standard_normal((N, K)).astype(np.float32))
    bias = vk.Array(gpu, data=rng.standard_normal(N).astype(np.float32))
    C, C2 = vk.Array(gpu, shape=(M, N)), vk.Array(gpu, shape=(M, N))
    cbf, cbft = vk.U32Array(gpu, shape=(M * N // 2,)), vk.U32Array(gpu, shape=(M * N // 2,))
    args = (md, 0, 1, M, N, K, 1.0, A._ptr, K, B._ptr, K, 0.0)
    tail = (N, bias._ptr, 1, cbf._ptr, None, None, 0, 0, None, None, cbft._ptr, M, None, 1)
    _lib.call("vkh_gemm_set_config", h, md, 0, 1, M, N, K, -1)
    _lib.call("vkh_gemm_bf16io_ext", h, *args, C._ptr, *tail)
    gpu.wait()
    before = np.asarray(C).copy()
    n = len(vk.gemm_configs(mode))
    ms = (ctypes.c_float * n)()
    best = ctypes.c_int(-1)
    _lib.call("vkh_gemm_autotune_ext", h, *args, C2._ptr, *tail, 2, ctypes.byref(best), ms)
    assert 0 <= best.value < n
    assert all(t > 0 for t in ms) and ms[best.value] == min(ms)
    idx, tuned = ctypes.c_int(), ctypes.c_int()
    _lib.call("vkh_gemm_get_config", h, md, 0, 1, M, N, K, ctypes.byref(idx), ctypes.byref(tuned))
    assert (idx.value, tuned.value) == (best.value, 1)
    _lib.call("vkh_gemm_bf16io_ext", h, *args, C2._ptr, *tail)
    gpu.wait()
    np.testing.assert_array_equal(np.asarray(C2).view(np.uint32), before.view(np.uint32))
    _lib.call("vkh_gemm_set_config", h, md, 0, 1, M, N, K, -1)


# Split k: few output tiles (BF16 modes: of 256 x 256 with k >= 4096, chunks
# of at least 1024; FP32: of 128 x 128 with k >= 1024, chunks of at least
# 256) run k in chunks as separate blocks, then sum the chunk partials in
# order.
SPLITK = [(300, 200, 8192), (300, 70, 5000), (64, 48, 4100), (4096, 256, 16384)]


@pytest.mark.parametrize("mode", ["fp32", "bf16", "bf16x3"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1)])
@pytest.mark.parametrize("M,N,K", SPLITK[:3])
def test_split_k_accuracy(gpu, mode, ta, tb, M, N, K):
    rng = np.random.default_rng(M + N + K + 3 * ta + tb)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    got = run_gemm(gpu, mode, ta, tb, A, B)
    ref, bound = ref_gemm(ta, tb, A, B)
    assert (np.abs(got - ref) / bound).max() <= TOL[mode]


def test_split_k_exact_and_epilogue(gpu):
    rng = np.random.default_rng(21)
    # integer operands: every chunk partial and their sum are exact
    A = rng.integers(-2, 3, (64, 8192)).astype(np.float32)
    B = rng.integers(-2, 3, (8192, 48)).astype(np.float32)
    for mode in ("fp32", "bf16", "bf16x3"):
        np.testing.assert_array_equal(run_gemm(gpu, mode, 0, 0, A, B), A @ B)
    # alpha, beta * C, bias and ReLU applied once, after the chunks are summed
    A = rng.standard_normal((300, 5000)).astype(np.float32)
    B = rng.standard_normal((70, 5000)).astype(np.float32)
    C0 = rng.standard_normal((300, 70)).astype(np.float32)
    bias = rng.standard_normal(70).astype(np.float32)
    ref, bound = ref_gemm(0, 1, A, B)
    for mode in ("fp32", "bf16", "bf16x3"):
        got = run_gemm(gpu, mode, 0, 1, A, B, C0=C0, alpha=0.5, beta=2.0, bias=bias)
        want = 0.5 * ref + 2.0 * C0 + bias
        assert (np.abs(got - want) <= 0.5 * TOL[mode] * bound + 1e-5 * (1 + np.abs(want))).all()
        got = run_gemm(gpu, mode, 0, 1, A, B, bias=bias, act="relu")
        want = np.maximum(ref + bias, 0)
        assert (np.abs(got - want) <= TOL[mode] * bound + 1e-5 * (1 + np.abs(want))).all()


def test_split_k_weight_gradient_shape(gpu):
    """dW = dY^T X of a 256 -> 4096 layer at batch 16384 (16 chunks)."""
    M, N, K = SPLITK[3]
    rng = np.random.default_rng(22)
    A = rng.standard_normal((K, M)).astype(np.float32)   # dY [batch, out]
    B = rng.standard_normal((K, N)).astype(np.float32)   # X [batch, in]
    rows = np.arange(0, M, 97)
    a64 = A[:, rows].T.astype(np.float64)
    ref, bound = a64 @ B.astype(np.float64), np.abs(a64) @ np.abs(B.astype(np.float64))
    for mode in ("fp32", "bf16", "bf16x3"):
        got = run_gemm(gpu, mode, 1, 0, A, B)[rows]
        assert (np.abs(got - ref) / bound).max() <= TOL[mode]


# Short k (K <= 16): FP32 vector kernel in every mode, epilogue included.
SMALLK = [(300, 260, 1), (257, 300, 16), (1000, 70, 10), (64, 4097, 3), (40, 40, 0)]


@pytest.mark.parametrize("mode", ["fp32", "bf16", "bf16x3"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", SMALLK)
def test_short_k(gpu, mode, ta, tb, M, N, K):
    rng = np.random.default_rng(M + 5 * N + K + ta + 2 * tb)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    got = run_gemm(gpu, mode, ta, tb, A, B, bias=bias, act="relu")
    ref, bound = ref_gemm(ta, tb, A, B)
    want = np.maximum(ref + bias, 0)
    # FP32 products summed in order: K roundings of |A||B| at most, plus the bias
    assert (np.abs(got - want) <= max(K, 1) * 2.0 ** -23 * (bound + np.abs(bias)) + 1e-30).all()


def test_short_k_aux_epilogue(gpu):
    """dX = dY W times ReLU'(h): the Dense input gradient of a 1-output layer."""
    rng = np.random.default_rng(30)
    B_, H = 512, 300
    dy = rng.standard_normal((B_, 1)).astype(np.float32)
    W = rng.standard_normal((1, H)).astype(np.float32)
    h = np.maximum(rng.standard_normal((B_, H)), 0).astype(np.float32)
    g_dy, g_W, g_h = vk.Array(gpu, data=dy), vk.Array(gpu, data=W), vk.Array(gpu, data=h)
    out = vk.Array(gpu, shape=(B_, H))
    for mode in ("fp32", "bf16", "bf16x3"):
        _lib.call("vkh_gemm_ex", gpu._handle, _lib.GEMM_MODE[mode], 0, 0, B_, H, 1, 1.0,
                  g_dy._ptr, 1, g_W._ptr, H, 0.0, out._ptr, H, None, 0, None,
                  g_h._ptr, H, _lib.ACT["relu"])
        out._set_job([g_dy, g_W, g_h])
        want = (dy * W) * (h > 0)
        np.testing.assert_array_equal(np.asarray(out), want.astype(np.float32))


def test_prepacked_bf16_operand(gpu):
    """vkh_gemm_bf16a with op(A) given as BF16 rows (M rounded up to 256,
    K % 64 == 0) gives the same bits as packing A, in the products the
    captured step hands activations through: plain, with bias and ReLU, and
    split k; other modes and layouts ignore the BF16 copy."""
    rng = np.random.default_rng(40)
    for (M, N, K, act) in ((300, 200, 128, "none"), (300, 260, 256, "relu"), (64, 48, 8192, "none")):
        a = rng.standard_normal((M, K)).astype(np.float32)
        b = rng.standard_normal((N, K)).astype(np.float32)
        bias = rng.standard_normal(N).astype(np.float32)
        gA, gB, gbias = vk.Array(gpu, data=a), vk.Array(gpu, data=b), vk.Array(gpu, data=bias)
        rows = -(-M // 256) * 256
        a_bf = vk.U32Array(gpu, shape=(rows * K // 2,))       # rows x K BF16 values
        _lib.call("vkh_to_bf16", gpu._handle, gA._ptr, a_bf._ptr, M * K)
        outs = {}
        for mode in ("bf16", "bf16x3", "fp32"):
            for pre in (False, True):
                C = vk.Array(gpu, shape=(M, N))
                c_bf = vk.U32Array(gpu, shape=((M * N + 1) // 2,))
                _lib.call("vkh_gemm_bf16a", gpu._handle, _lib.GEMM_MODE[mode], 0, 1, M, N, K, 1.0,
                          gA._ptr, K, gB._ptr, K, 0.0, C._ptr, N, gbias._ptr, _lib.ACT[act],
                          c_bf._ptr, None, 0, 0, a_bf._ptr if pre else None)
                C._set_job([gA, gB, a_bf])
                outs[(mode, pre)] = (np.asarray(C).copy(), np.asarray(c_bf).copy())
            np.testing.assert_array_equal(outs[(mode, True)][0], outs[(mode, False)][0], err_msg=mode)
            # the BF16 copy of C is the BF16 rounding of C
            conv = vk.U32Array(gpu, shape=((M * N + 1) // 2,))
            gC = vk.Array(gpu, data=outs[(mode, True)][0])
            _lib.call("vkh_to_bf16", gpu._handle, gC._ptr, conv._ptr, M * N)
            conv._set_job([gC])
            np.testing.assert_array_equal(np.asarray(conv)[: M * N // 2],
                                          outs[(mode, True)][1][: M * N // 2], err_msg=mode)




def test_to_bf16_rounding(gpu):
    """vkh_to_bf16 on ties, denormals, values that would round to zero, inf
    and NaN: the library's rounding as the NumPy statement gives it."""
    rng = np.random.default_rng(41)
    special = np.array([0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 2 ** -133, 2 ** -134, 1e-45, -1e-45,
                        -3e-41, 2 ** -126, np.inf, -np.inf, 3.4e38, -3.4e38], dtype=np.float32)
    x = np.concatenate([special, rng.standard_normal(4096).astype(np.float32),
                        (rng.standard_normal(4096) * 1e-40).astype(np.float32)])
    X = vk.Array(gpu, data=x)
    out = vk.U32Array(gpu, shape=((x.size + 1) // 2,))
    _lib.call("vkh_to_bf16", gpu._handle, X._ptr, out._ptr, x.size)
    out._set_job([X])
    got = np.asarray(out).view(np.uint16)[: x.size]
    np.testing.assert_array_equal(got, bf16_bits(x))
    nan = np.array([np.nan, -np.nan], dtype=np.float32)
    Xn = vk.Array(gpu, data=nan)
    on = vk.U32Array(gpu, shape=(1,))
    _lib.call("vkh_to_bf16", gpu._handle, Xn._ptr, on._ptr, 2)
    on._set_job([Xn])
    b = np.asarray(on).view(np.uint16)
    assert all((int(v) & 0x7F80) == 0x7F80 and (int(v) & 0x7F) != 0 for v in b)


@pytest.mark.parametrize("mode", ["fp32", "bf16"])
@pytest.mark.parametrize("M,N,K", [(300, 200, 96), (300, 198, 96), (300, 8, 96), (300, 200, 8),
                                   (512, 512, 8192)])
def test_gemm_ext_bf16_aux_and_no_fp32_c(gpu, mode, M, N, K):
    """vkh_gemm_bf16io_ext: ReLU' read from the BF16 copy of the activation
    gives C bit for bit as vkh_gemm_ex with the FP32 activation, on every
    path (matrix cores, ragged rows, skinny, short k, split k), including
    activations that are -0, denormal or too small for BF16's range; with
    store_c = 0 C is left as it was (also with C NULL) and the BF16 copy is
    the library's rounding of the values."""
    h = gpu._handle
    md = _lib.GEMM_MODE[mode]
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    aux = np.maximum(rng.standard_normal((M, N)), 0).astype(np.float32)
    flat = aux.reshape(-1)
    idx = rng.choice(flat.size, 64, replace=False)
    flat[idx[:16]] = -0.0
    flat[idx[16:32]] = 1e-45
    flat[idx[32:48]] = 3e-41
    flat[idx[48:]] = 2 ** -126
    gA, gB, gX = vk.Array(gpu, data=A), vk.Array(gpu, data=B), vk.Array(gpu, data=aux)
    x16 = vk.U32Array(gpu, shape=((M * N + 1) // 2,))
    _lib.call("vkh_to_bf16", h, gX._ptr, x16._ptr, M * N)
    ref = vk.Array(gpu, shape=(M, N))
    _lib.call("vkh_gemm_ex", h, md, 0, 0, M, N, K, 1.0, gA._ptr, K, gB._ptr, N, 0.0, ref._ptr, N,
              None, 0, None, gX._ptr, N, 1)
    C = vk.Array(gpu, shape=(M, N))
    _lib.call("vkh_gemm_bf16io_ext", h, md, 0, 0, M, N, K, 1.0, gA._ptr, K, gB._ptr, N, 0.0, C._ptr, N,
              None, 0, None, None, x16._ptr, N, 1, None, None, None, 0, None, 1)
    gpu.wait()
    want = np.asarray(ref)
    np.testing.assert_array_equal(np.asarray(C).view(np.uint32), want.view(np.uint32))
    for c_arg in ("keep", None):
        C2 = vk.Array(gpu, data=np.full((M, N), 7.0, dtype=np.float32))
        cbf = vk.U32Array(gpu, shape=((M * N + 1) // 2,))
        _lib.call("vkh_gemm_bf16io_ext", h, md, 0, 0, M, N, K, 1.0, gA._ptr, K, gB._ptr, N, 0.0,
                  C2._ptr if c_arg else None, N, None, 0, cbf._ptr, None, x16._ptr, N, 1, None, None, None, 0,
                  None, 0)
        for o in (C2, cbf):
            o._set_job([gA, gB, x16])
        gpu.wait()
        assert np.all(np.asarray(C2) == 7.0)
        np.testing.assert_array_equal(np.asarray(cbf).view(np.uint16)[: M * N].reshape(M, N), bf16_bits(want))
    with pytest.raises(ValueError):  # a BF16 aux is for ReLU' only
        _lib.call("vkh_gemm_bf16io_ext", h, md, 0, 0, M, N, K, 1.0, gA._ptr, K, gB._ptr, N, 0.0, C._ptr, N,
                  None, 0, None, None, x16._ptr, N, 2, None, None, None, 0, None, 1)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_bf16io_transposed_output_and_packed_b(gpu, ta, tb):
    """vkh_gemm_bf16io: c_bf16t[n * ldt + m] is the BF16 rounding of C[m, n]
    on every path (matrix cores, split k, skinny, short k), for a 16-byte
    aligned and an unaligned ldt, and nothing else of c_bf16t is written;
    op(A) and op(B)^T given as packed BF16 rows leave C's bits unchanged."""
    rng = np.random.default_rng(41 + 2 * ta + tb)
    shapes = ((300, 260, 256, "relu"), (256, 192, 8192, "none"), (8, 300, 128, "none"),
              (300, 8, 64, "none"), (300, 260, 8, "none"), (130, 70, 64, "sigmoid"))
    for M, N, K, act in shapes:
        A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
        Bm = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
        bias = rng.standard_normal(N).astype(np.float32)
        gA, gB, gbias = vk.Array(gpu, data=A), vk.Array(gpu, data=Bm), vk.Array(gpu, data=bias)
        h, mode = gpu._handle, _lib.GEMM_MODE["bf16"]

        def run(a_pre=None, b_pre=None, ct=None, ldt=0):
            C = vk.Array(gpu, shape=(M, N))
            _lib.call("vkh_gemm_bf16io", h, mode, ta, tb, M, N, K, 1.0, gA._ptr, A.shape[1], gB._ptr,
                      Bm.shape[1], 0.0, C._ptr, N, gbias._ptr, _lib.ACT[act], None, None, 0, 0,
                      a_pre._ptr if a_pre else None, b_pre._ptr if b_pre else None,
                      ct._ptr if ct else None, ldt)
            C._set_job([gA, gB] + [t for t in (a_pre, b_pre, ct) if t])
            if ct:
                ct._set_job([])   # written by the product: drop the host mirror
            return np.asarray(C).copy()

        C_ref = run()
        want_t = bf16_bits(C_ref).T
        for ldt in ((M + 7) // 8 * 8 + 8, M + 5):
            ct = vk.U32Array(gpu, data=np.full((N * ldt + 1) // 2, 0xFFFFFFFF, np.uint32))
            np.testing.assert_array_equal(run(ct=ct, ldt=ldt), C_ref, err_msg=(M, N, K))
            got = np.asarray(ct).view(np.uint16)[: N * ldt].reshape(N, ldt)
            np.testing.assert_array_equal(got[:, :M], want_t, err_msg=(M, N, K, ldt))
            assert (got[:, M:] == 0xFFFF).all(), (M, N, K, ldt)
        # packed operands: op(A) rows and op(B)^T rows, 256-row padded, zero tails
        pre = []
        for rowsmat in (A.T if ta else A, Bm if tb else Bm.T):
            R = rowsmat.shape[0]
            buf = np.zeros(-(-R // 256) * 256 * K, np.uint16)
            buf[: R * K] = bf16_bits(rowsmat).ravel()
            pre.append(vk.U32Array(gpu, data=buf.view(np.uint32)))
        np.testing.assert_array_equal(run(a_pre=pre[0]), C_ref, err_msg=(M, N, K))
        np.testing.assert_array_equal(run(b_pre=pre[1]), C_ref, err_msg=(M, N, K))
        np.testing.assert_array_equal(run(a_pre=pre[0], b_pre=pre[1]), C_ref, err_msg=(M, N, K))
    # misaligned hand-off buffers are refused on the host, before any launch
    # (the offsets stay inside `big`, 256 KiB, whatever the product would touch)
    big = vk.U32Array(gpu, shape=(256 * 256,))
    for a_p, b_p, c_p in ((big._ptr + 2, None, None), (None, big._ptr + 8, None),
                          (None, None, big._ptr + 1)):
        C = vk.Array(gpu, shape=(M, N))
        with pytest.raises(ValueError):
            _lib.call("vkh_gemm_bf16io", gpu._handle, _lib.GEMM_MODE["bf16"], ta, tb, M, N, K, 1.0,
                      gA._ptr, A.shape[1], gB._ptr, Bm.shape[1], 0.0, C._ptr, N, None, 0, None, None, 0,
                      0, a_p, b_p, c_p, M)
    gpu.wait()


@pytest.mark.parametrize("M,N,K", [(300, 260, 256), (1024, 512, 256)])
def test_every_bf16_configuration_with_bf16_copies(gpu, M, N, K):
    """Every BF16 kernel configuration with the forward epilogue (bias, ReLU)
    and both BF16 copies of C, row-major and transposed (each configuration's
    own store path for C^T): C bit for bit the same in every configuration,
    each copy the library's rounding of C, and C within the FP32 accumulation
    bound of relu(A W^T + b) computed in FP64 from the BF16-rounded operands."""
    rng = np.random.default_rng(M + 3 * N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    W = rng.standard_normal((N, K)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    gA, gW, gb = vk.Array(gpu, data=A), vk.Array(gpu, data=W), vk.Array(gpu, data=bias)
    h, md = gpu._handle, _lib.GEMM_MODE["bf16"]
    ldt = (M + 7) // 8 * 8
    cfgs = vk.gemm_configs("bf16")
    outs = []
    try:
        for i in range(len(cfgs)):
            _lib.call("vkh_gemm_set_config", h, md, 0, 1, M, N, K, i)
            C = vk.Array(gpu, shape=(M, N))
            cbf = vk.U32Array(gpu, shape=((M * N + 1) // 2,))
            ct = vk.U32Array(gpu, shape=((N * ldt + 1) // 2,))
            _lib.call("vkh_gemm_bf16io", h, md, 0, 1, M, N, K, 1.0, gA._ptr, K, gW._ptr, K, 0.0, C._ptr, N,
                      gb._ptr, _lib.ACT["relu"], cbf._ptr, None, 0, 0, None, None, ct._ptr, ldt)
            C._set_job([gA, gW, gb])
            cbf._set_job([])   # written by the product: drop the host mirrors
            ct._set_job([])
            c = np.asarray(C).copy()
            outs.append(c)
            want = bf16_bits(c)
            np.testing.assert_array_equal(np.asarray(cbf).view(np.uint16)[: M * N].reshape(M, N), want,
                                          err_msg=f"row copy, configuration {i} {cfgs[i]}")
            got_t = np.asarray(ct).view(np.uint16)[: N * ldt].reshape(N, ldt)[:, :M]
            np.testing.assert_array_equal(got_t, want.T, err_msg=f"transposed copy, configuration {i} {cfgs[i]}")
    finally:
        _lib.call("vkh_gemm_set_config", h, md, 0, 1, M, N, K, -1)
    for i, o in enumerate(outs[1:], 1):
        np.testing.assert_array_equal(o.view(np.uint32), outs[0].view(np.uint32),
                                      err_msg=f"configuration {i} {cfgs[i]}")
    a16 = (bf16_bits(A).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    w16 = (bf16_bits(W).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    ref = np.maximum(a16 @ w16.T + bias.astype(np.float64), 0.0)
    bound = np.abs(a16) @ np.abs(w16).T + np.abs(bias.astype(np.float64))
    assert (np.abs(outs[0] - ref) / bound).max() <= (K + 2) * 2.0 ** -24
