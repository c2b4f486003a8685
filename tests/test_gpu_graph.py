"""The captured training step (nn.GraphSequence) against the eager
Sequence.train path: identical weights in FP32 after every replay (with eager
work and synchronisations between replays), and a BF16 run that tracks the
FP32 loss."""
import ctypes
import gc

import numpy as np
import pytest
from bf16ref import bf16_bits

import vulkpy_amd as vk
from vulkpy_amd import _lib, nn

pytestmark = pytest.mark.gpu


def make_net(gpu, dims, acts, seed, reg=None, opt="adam", precision=None):
    layers = []
    for k, (i, o) in enumerate(dims):
        if opt == "adam":
            wo, bo = nn.Adam(gpu, lr=1e-3), nn.Adam(gpu, lr=1e-3)
        elif opt == "adagrad":
            wo, bo = nn.AdaGrad(gpu, lr=1e-2, tau=0.1), nn.AdaGrad(gpu, lr=1e-2, tau=0.1)
        else:
            wo, bo = nn.SGD(0.05), nn.SGD(0.05)
        layers.append(nn.Dense(gpu, i, o, w_init=nn.HeNormal(gpu, i, seed=seed + k),
                               w_opt=wo, b_opt=bo, w_reg=reg() if reg else None,
                               b_reg=reg() if reg and opt != "sgd" else None,
                               precision=precision))
        if acts[k] == "relu":
            layers.append(nn.ReLU())
        elif acts[k] == "sigmoid":
            layers.append(nn.Sigmoid())
    return layers


def dense(layers):
    return [L for L in layers if isinstance(L, nn.Dense)]


def weights(layers):
    return [np.asarray(p).copy() for L in layers if isinstance(L, nn.Dense) for p in (L.w.value, L.b.value)]


CASES = [
    # (dims, acts, loss, regularizer, optimizer)
    ([(20, 32), (32, 24), (24, 10)], ["relu", "sigmoid", None], "sce", None, "adam"),
    ([(16, 64), (64, 64), (64, 1)], ["relu", "relu", None], "huber", lambda: nn.Ridge(1e-4), "adam"),
    ([(8, 16), (16, 3)], ["sigmoid", None], "mse", lambda: nn.Lasso(1e-3), "sgd"),
    ([(12, 40), (40, 4)], ["relu", None], "sce", lambda: nn.Elastic(1e-3, 2e-3), "adagrad"),
    ([(12, 40), (40, 40), (40, 2)], ["relu", "relu", None], "huber", lambda: nn.Elastic(5e-4, 1e-3), "adam"),
    # ten Adam parameters: two multi-parameter update launches
    ([(10, 24), (24, 24), (24, 16), (16, 16), (16, 3)], ["relu", "sigmoid", "relu", "relu", None], "mse",
     lambda: nn.Ridge(1e-3), "adam"),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_graph_fp32_matches_eager(gpu, case):
    dims, acts, loss_name, reg, opt = CASES[case]
    B = 96
    rng = np.random.default_rng(case)
    x = rng.standard_normal((B, dims[0][0])).astype(np.float32)
    O = dims[-1][1]
    if loss_name == "sce":
        y = np.identity(O, dtype=np.float32)[rng.integers(0, O, B)]
    else:
        y = rng.standard_normal((B, O)).astype(np.float32)
    mk_loss = {"sce": nn.SoftmaxCrossEntropyLoss, "huber": nn.HuberLoss, "mse": nn.MSELoss}[loss_name]
    eager_layers = make_net(gpu, dims, acts, 100, reg, opt)
    graph_layers = make_net(gpu, dims, acts, 100, reg, opt)
    eager = nn.Sequence(eager_layers, mk_loss(), regularize=True)
    graph = nn.GraphSequence(graph_layers, mk_loss(), precision="fp32", regularize=True)
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    for step in range(4):
        _, le = eager.train(X, Y)
        _, lg = graph.train(X, Y)
        gpu.wait()
        assert np.asarray(le).item() == np.asarray(lg).item(), step
        for a, b in zip(weights(eager_layers), weights(graph_layers)):
            np.testing.assert_array_equal(a, b)
        # eager work between replays
        (X * 2.0).sum()
        gpu.wait()


def test_graph_returns_fresh_arrays(gpu):
    """Each train call returns its own prediction and loss arrays (as the
    reference's Sequence.train does), so results kept across steps are not
    overwritten by later replays."""
    dims, acts = [(16, 64), (64, 64), (64, 1)], ["relu", "relu", None]
    rng = np.random.default_rng(3)
    x = rng.standard_normal((64, 16)).astype(np.float32)
    y = rng.standard_normal((64, 1)).astype(np.float32)
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    eager = nn.Sequence(make_net(gpu, dims, acts, 21, lambda: nn.Ridge(1e-4)), nn.HuberLoss(),
                        regularize=True)
    graph = nn.GraphSequence(make_net(gpu, dims, acts, 21, lambda: nn.Ridge(1e-4)), nn.HuberLoss(),
                             precision="fp32")
    kept, want = [], []
    for _ in range(5):
        pe, le = eager.train(X, Y)
        want.append((np.asarray(pe).copy(), np.asarray(le).item()))
        kept.append(graph.train(X, Y))
    assert len({id(a) for pair in kept for a in pair}) == 10
    for (pg, lg), (pe, le) in zip(kept, want):
        np.testing.assert_array_equal(np.asarray(pg), pe)
        assert np.asarray(lg).item() == le
    assert want[0][1] != want[-1][1]


def test_graph_bf16_tracks_fp32(gpu):
    dims, acts = [(32, 128), (128, 128), (128, 1)], ["relu", "relu", None]
    B = 512
    rng = np.random.default_rng(7)
    x = rng.standard_normal((B, 32)).astype(np.float32)
    P = rng.standard_normal((32, 1)).astype(np.float32)
    y = (np.sin(x @ P) + 0.1 * rng.standard_normal((B, 1))).astype(np.float32)
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    f32 = nn.GraphSequence(make_net(gpu, dims, acts, 5, lambda: nn.Ridge(1e-4)), nn.HuberLoss(),
                           precision="fp32")
    b16 = nn.GraphSequence(make_net(gpu, dims, acts, 5, lambda: nn.Ridge(1e-4)), nn.HuberLoss(),
                           precision="bf16")
    l32 = [float(np.asarray(f32.train(X, Y)[1])) for _ in range(60)]
    l16 = [float(np.asarray(b16.train(X, Y)[1])) for _ in range(60)]
    assert l32[-1] < 0.7 * l32[0]
    assert abs(l16[-1] - l32[-1]) <= 0.05 * l32[-1] + 1e-3, (l16[-1], l32[-1])


@pytest.mark.parametrize("precision", ["bf16", "bf16x3"])
def test_graph_packed_matches_eager_with_interleaved_gemms(gpu, precision):
    """Captured BF16 GEMMs pack into their graph's own region: eager BF16
    products between replays (one of them larger than any product of the
    step, which regrows the eager packing area) leave the replays identical
    to the eager step with the same kernels."""
    dims, acts = [(24, 96), (96, 80), (80, 5)], ["relu", "sigmoid", None]
    B = 160
    rng = np.random.default_rng(11)
    x = rng.standard_normal((B, 24)).astype(np.float32)
    y = np.identity(5, dtype=np.float32)[rng.integers(0, 5, B)]
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    eager_layers = make_net(gpu, dims, acts, 40, lambda: nn.Ridge(1e-3), precision=precision)
    graph_layers = make_net(gpu, dims, acts, 40, lambda: nn.Ridge(1e-3), precision=precision)
    eager = nn.Sequence(eager_layers, nn.SoftmaxCrossEntropyLoss(), regularize=True)
    graph = nn.GraphSequence(graph_layers, nn.SoftmaxCrossEntropyLoss(), precision=precision,
                             regularize=True)
    for step in range(4):
        _, le = eager.train(X, Y)
        _, lg = graph.train(X, Y)
        gpu.wait()
        assert np.asarray(le).item() == np.asarray(lg).item(), step
        for a, b in zip(weights(eager_layers), weights(graph_layers)):
            np.testing.assert_array_equal(a, b)
        n = 300 + 150 * step  # grows beyond every product of the step
        sq = vk.Array(gpu, data=rng.standard_normal((n, n)).astype(np.float32))
        sq.matmul(sq, precision=precision)
        gpu.wait()


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp32"])
def test_graph_split_k_matches_eager(gpu, precision):
    """At batch 8192 the weight gradients of the 64 -> 256 -> 64 layers split
    k into chunks (8 in the BF16 modes, 32 in FP32); the captured step (chunk
    partials in its packing region) stays identical to the eager step."""
    dims, acts = [(64, 256), (256, 64), (64, 4)], ["relu", "sigmoid", None]
    B = 8192
    rng = np.random.default_rng(13)
    x = rng.standard_normal((B, 64)).astype(np.float32)
    y = rng.standard_normal((B, 4)).astype(np.float32)
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    eager_layers = make_net(gpu, dims, acts, 70, lambda: nn.Ridge(1e-4), precision=precision)
    graph_layers = make_net(gpu, dims, acts, 70, lambda: nn.Ridge(1e-4), precision=precision)
    eager = nn.Sequence(eager_layers, nn.MSELoss(), regularize=True)
    graph = nn.GraphSequence(graph_layers, nn.MSELoss(), precision=precision, regularize=True)
    for step in range(3):
        _, lg = graph.train(X, Y)
        _, le = eager.train(X, Y)
        gpu.wait()
        assert np.asarray(le).item() == np.asarray(lg).item(), step
        for a, b in zip(weights(eager_layers), weights(graph_layers)):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp32"])
def test_graph_uncaptured_matches_captured(gpu, precision):
    """capture=False issues the step's kernels one launch at a time: the
    losses and weights equal the captured step's after every step, BF16
    hand-off buffers included."""
    dims, acts = [(64, 320), (320, 192), (192, 3)], ["relu", "relu", None]
    B = 512
    rng = np.random.default_rng(17)
    x = rng.standard_normal((B, 64)).astype(np.float32)
    y = rng.standard_normal((B, 3)).astype(np.float32)
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    cap_layers = make_net(gpu, dims, acts, 90, lambda: nn.Ridge(1e-4), precision=precision)
    unc_layers = make_net(gpu, dims, acts, 90, lambda: nn.Ridge(1e-4), precision=precision)
    cap = nn.GraphSequence(cap_layers, nn.HuberLoss(), precision=precision, regularize=True)
    unc = nn.GraphSequence(unc_layers, nn.HuberLoss(), precision=precision, regularize=True,
                           capture=False)
    for step in range(3):
        _, lc = cap.train(X, Y)
        _, lu = unc.train(X, Y)
        gpu.wait()
        assert unc._plan.graph is None
        if precision == "bf16":
            # h0 and dY1 go to the next products as BF16 copies, and h0^T,
            # dY0^T to the weight gradients as transposed BF16 copies
            p = unc._plan
            assert p.hs_bf[0] is not None and p.ds_bf[1] is not None
            assert p.hs_bft[0] is not None and p.ds_bft[0] is not None
        assert np.asarray(lc).item() == np.asarray(lu).item(), step
        for a, b in zip(weights(cap_layers), weights(unc_layers)):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision,hidden_act", [("fp32", "relu"), ("fp32", "sigmoid"), ("fp32", None),
                                                  ("bf16", "relu"), ("bf16x3", "sigmoid")])
def test_graph_fused_last_layer_matches_eager(gpu, precision, hidden_act):
    """A one-output last layer at batch 1024: its input gradient (and BF16
    copies), weight gradient and the hidden layer's bias gradient come from
    one kernel (vkh_dense1_backward) with the eager step's bits."""
    dims, acts = [(32, 256), (256, 192), (192, 1)], ["relu", hidden_act, None]
    B = 1024
    rng = np.random.default_rng(23)
    x = rng.standard_normal((B, 32)).astype(np.float32)
    y = rng.standard_normal((B, 1)).astype(np.float32)
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    eager_layers = make_net(gpu, dims, acts, 50, lambda: nn.Ridge(1e-4), precision=precision)
    graph_layers = make_net(gpu, dims, acts, 50, lambda: nn.Ridge(1e-4), precision=precision)
    eager = nn.Sequence(eager_layers, nn.HuberLoss(), regularize=True)
    graph = nn.GraphSequence(graph_layers, nn.HuberLoss(), precision=precision, regularize=True)
    for step in range(3):
        _, le = eager.train(X, Y)
        _, lg = graph.train(X, Y)
        gpu.wait()
        p = graph._plan
        assert p.fuse_last
        if precision == "bf16":
            assert p.ds_bf[1] is not None and p.ds_bft[1] is not None
        assert np.asarray(le).item() == np.asarray(lg).item(), step
        for a, b in zip(weights(eager_layers), weights(graph_layers)):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("B,N", [(512, 96), (2048, 4096)])
def test_dense1_backward_matches_separate_steps(gpu, B, N):
    """vkh_dense1_backward against vkh_gemm (input and weight gradients) and
    vkh_bias_grad, bit for bit, BF16 copies included, on a 512 x 96 layer
    and a 2048 x 4096 one (whose blocks take two row tiles each)."""
    h = gpu._handle
    rng = np.random.default_rng(29)
    dy = vk.Array(gpu, data=rng.standard_normal((B, 1)).astype(np.float32))
    w = vk.Array(gpu, data=rng.standard_normal((1, N)).astype(np.float32))
    hid = vk.Array(gpu, data=np.maximum(rng.standard_normal((B, N)), 0).astype(np.float32))
    ok = ctypes.c_int(0)
    _lib.call("vkh_dense1_backward_supported", h, B, N, ctypes.byref(ok))
    assert ok.value == 1
    dx, dw, db = (vk.Array(gpu, shape=s) for s in ((B, N), (1, N), (N,)))
    dxbf = vk.U32Array(gpu, shape=(B * N // 2,))
    dxbft = vk.U32Array(gpu, shape=(B * N // 2,))
    _lib.call("vkh_dense1_backward", h, dy._ptr, w._ptr, hid._ptr, 1, B, N, dx._ptr, dxbf._ptr,
              dxbft._ptr, B, dw._ptr, db._ptr)
    rx, rw, rb = (vk.Array(gpu, shape=s) for s in ((B, N), (1, N), (N,)))
    _lib.call("vkh_gemm_ex", h, 0, 0, 0, B, N, 1, 1.0, dy._ptr, 1, w._ptr, N, 0.0, rx._ptr, N,
              None, 0, None, hid._ptr, N, 1)
    _lib.call("vkh_gemm", h, 0, 1, 0, 1, N, B, 1.0, dy._ptr, 1, hid._ptr, N, 0.0, rw._ptr, N,
              None, 0, None)
    _lib.call("vkh_bias_grad", h, rx._ptr, rb._ptr, B, N, 0.0)
    gpu.wait()
    got_x = np.asarray(dx)
    np.testing.assert_array_equal(got_x.view(np.uint32), np.asarray(rx).view(np.uint32))
    np.testing.assert_array_equal(np.asarray(dw).view(np.uint32), np.asarray(rw).view(np.uint32))
    np.testing.assert_array_equal(np.asarray(db).view(np.uint32), np.asarray(rb).view(np.uint32))
    # BF16 copies: the library's rounding of the FP32 values
    bf = bf16_bits(got_x)
    np.testing.assert_array_equal(np.asarray(dxbf).view(np.uint16).reshape(B, N), bf)
    np.testing.assert_array_equal(np.asarray(dxbft).view(np.uint16).reshape(N, B), bf.T)
    # without the FP32 dx: the same copies, weight and bias gradients
    dw2, db2 = vk.Array(gpu, shape=(1, N)), vk.Array(gpu, shape=(N,))
    dxbf2, dxbft2 = vk.U32Array(gpu, shape=(B * N // 2,)), vk.U32Array(gpu, shape=(B * N // 2,))
    _lib.call("vkh_dense1_backward", h, dy._ptr, w._ptr, hid._ptr, 1, B, N, None, dxbf2._ptr,
              dxbft2._ptr, B, dw2._ptr, db2._ptr)
    gpu.wait()
    for a, b in ((dw2, dw), (db2, db), (dxbf2, dxbf), (dxbft2, dxbft)):
        np.testing.assert_array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.mark.parametrize("mode", ["fp32", "bf16"])
def test_gemm_colsum_matches_separate_steps(gpu, mode):
    """vkh_gemm_bf16io_colsum: C as vkh_gemm_ex gives it and C's column sums
    as vkh_bias_grad gives them, bit for bit; with store_c = 0 the FP32 C is
    left as it was.  2048 x 32768 x 64 with a ReLU' factor: the shape's bias
    gradient sums 128-row chunks, as the epilogue does."""
    M, N, K = 2048, 32768, 64
    h = gpu._handle
    md = _lib.GEMM_MODE[mode]
    ok = ctypes.c_int(0)
    _lib.call("vkh_gemm_colsum_supported", h, md, 0, 0, M, N, K, ctypes.byref(ok))
    assert ok.value == 1
    rng = np.random.default_rng(37)
    A = vk.Array(gpu, data=rng.standard_normal((M, K)).astype(np.float32))
    B = vk.Array(gpu, data=rng.standard_normal((K, N)).astype(np.float32))
    aux = vk.Array(gpu, data=np.maximum(rng.standard_normal((M, N)), 0).astype(np.float32))
    ref, db_ref = vk.Array(gpu, shape=(M, N)), vk.Array(gpu, shape=(N,))
    _lib.call("vkh_gemm_ex", h, md, 0, 0, M, N, K, 1.0, A._ptr, K, B._ptr, N, 0.0, ref._ptr, N,
              None, 0, None, aux._ptr, N, 1)
    _lib.call("vkh_bias_grad", h, ref._ptr, db_ref._ptr, M, N, 0.0)
    for store in (1, 0):
        C = vk.Array(gpu, data=np.full((M, N), 7.0, dtype=np.float32))
        db = vk.Array(gpu, shape=(N,))
        cbft = vk.U32Array(gpu, shape=(M * N // 2,))
        _lib.call("vkh_gemm_bf16io_colsum", h, md, 0, 0, M, N, K, 1.0, A._ptr, K, B._ptr, N, 0.0,
                  C._ptr, N, None, 0, None, aux._ptr, N, 1, None, None, cbft._ptr, M, db._ptr, store)
        for out in (C, db, cbft):
            out._set_job([A, B, aux])   # written on the device: the host mirrors are stale
        gpu.wait()
        np.testing.assert_array_equal(np.asarray(db).view(np.uint32), np.asarray(db_ref).view(np.uint32))
        got = np.asarray(C)
        if store:
            np.testing.assert_array_equal(got.view(np.uint32), np.asarray(ref).view(np.uint32))
        else:
            assert np.all(got == 7.0)
        # the transposed BF16 copy: the library's rounding of the FP32 values
        bf = bf16_bits(np.asarray(ref))
        np.testing.assert_array_equal(np.asarray(cbft).view(np.uint16).reshape(N, M), bf.T)
        del C, db, cbft, got


def test_graph_sees_inputs_and_weights_changed_between_steps(gpu):
    """The step skips copying an input that is the array it copied last time
    and unchanged: changing the input in place, or weights outside the step,
    between replays must reach the next step (eager and captured stay equal)."""
    dims, acts = [(64, 256), (256, 128), (128, 1)], ["relu", "relu", None]
    B = 512
    rng = np.random.default_rng(43)
    X = vk.Array(gpu, data=rng.standard_normal((B, 64)).astype(np.float32))
    Y = vk.Array(gpu, data=rng.standard_normal((B, 1)).astype(np.float32))
    eager_layers = make_net(gpu, dims, acts, 80, lambda: nn.Ridge(1e-4), precision="bf16")
    graph_layers = make_net(gpu, dims, acts, 80, lambda: nn.Ridge(1e-4), precision="bf16")
    eager = nn.Sequence(eager_layers, nn.HuberLoss(), regularize=True)
    graph = nn.GraphSequence(graph_layers, nn.HuberLoss(), precision="bf16", regularize=True)
    for step in range(5):
        _, le = eager.train(X, Y)
        _, lg = graph.train(X, Y)
        gpu.wait()
        assert np.asarray(le).item() == np.asarray(lg).item(), step
        for a, b in zip(weights(eager_layers), weights(graph_layers)):
            np.testing.assert_array_equal(a, b)
        if step == 1:
            X *= 1.5                                  # same array, changed in place
        if step == 2:
            for layers in (eager_layers, graph_layers):
                dense(layers)[0].w.value *= 0.75      # weights changed outside the step
        if step == 3:
            for layers in (eager_layers, graph_layers):
                dense(layers)[1].w.value[3, 7] = -0.25  # through the host mirror
                dense(layers)[2].w.value[0, 0] = 0.5


def test_graph_capture_survives_finalizers(gpu, monkeypatch):
    """Objects collected while the step is captured (an array whose upload
    may still be queued, with a pinned mirror, and a generator) release
    their resources after the capture; the graph stays valid and matches
    the eager step."""
    from vulkpy_amd.nn import graph as graph_mod
    big = vk.Array(gpu, data=np.ones(1 << 22, dtype=np.float32))  # pinned mirror, chunked upload
    assert big._upload_pending
    doomed = [big, vk.random.Xoshiro128pp(gpu, seed=3)]
    del big
    orig = graph_mod._Plan._enqueue

    def enqueue(self):
        if doomed and vk.vkarray._captures:
            doomed.clear()
            gc.collect()                      # their finalizers run inside the capture
            assert vk.vkarray._deferred
        orig(self)

    monkeypatch.setattr(graph_mod._Plan, "_enqueue", enqueue)
    dims, acts = [(12, 40), (40, 4)], ["relu", None]
    B = 96
    rng = np.random.default_rng(31)
    X = vk.Array(gpu, data=rng.standard_normal((B, 12)).astype(np.float32))
    Y = vk.Array(gpu, data=np.identity(4, dtype=np.float32)[rng.integers(0, 4, B)])
    eager_layers = make_net(gpu, dims, acts, 60, opt="adagrad")
    graph_layers = make_net(gpu, dims, acts, 60, opt="adagrad")
    eager = nn.Sequence(eager_layers, nn.SoftmaxCrossEntropyLoss())
    graph = nn.GraphSequence(graph_layers, nn.SoftmaxCrossEntropyLoss(), precision="fp32")
    for step in range(3):
        _, le = eager.train(X, Y)
        _, lg = graph.train(X, Y)
        gpu.wait()
        assert graph._plan.graph is not None
        assert not doomed and not vk.vkarray._deferred
        assert np.asarray(le).item() == np.asarray(lg).item(), step
        for a, b in zip(weights(eager_layers), weights(graph_layers)):
            np.testing.assert_array_equal(a, b)


def test_graph_rejects_unsupported(gpu):
    with pytest.raises(ValueError):
        nn.GraphSequence([nn.ReLU()], nn.MSELoss())
    net = nn.GraphSequence(make_net(gpu, [(4, 3)], [None], 1), nn.MSELoss(), precision="fp32")
    with pytest.raises(ValueError):
        net.train(vk.Array(gpu, shape=(8, 5)), vk.Array(gpu, shape=(8, 3)))


def test_graph_autotuned_matches_eager(gpu):
    """autotune=True picks GEMM configurations per product shape (they never
    change results) and reduction configurations per bias-gradient shape
    (kept on the context, so the eager step uses them too): results stay
    identical to the eager step.  The graph's first step, which tunes, runs
    before the eager one."""
    dims, acts = [(64, 256), (256, 256), (256, 4)], ["relu", "relu", None]
    B = 512
    rng = np.random.default_rng(12)
    x = rng.standard_normal((B, 64)).astype(np.float32)
    y = rng.standard_normal((B, 4)).astype(np.float32)
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    eager_layers = make_net(gpu, dims, acts, 60, lambda: nn.Ridge(1e-4), precision="bf16")
    graph_layers = make_net(gpu, dims, acts, 60, lambda: nn.Ridge(1e-4), precision="bf16")
    eager = nn.Sequence(eager_layers, nn.HuberLoss(), regularize=True)
    graph = nn.GraphSequence(graph_layers, nn.HuberLoss(), precision="bf16", regularize=True,
                             autotune=True)
    try:
        for step in range(3):
            _, lg = graph.train(X, Y)
            _, le = eager.train(X, Y)
            gpu.wait()
            assert np.asarray(le).item() == np.asarray(lg).item(), step
            for a, b in zip(weights(eager_layers), weights(graph_layers)):
                np.testing.assert_array_equal(a, b)
        # the step's reductions were tuned too (bias gradients: [B, 256], [B, 4])
        tuned = ctypes.c_int()
        idx = ctypes.c_int()
        _lib.call("vkh_reduce_get_config", gpu._handle, 1, 1, B, 256, ctypes.byref(idx),
                  ctypes.byref(tuned))
        assert tuned.value == 1
    finally:
        # reduction choices stay on the context: restore the defaults
        for d in (1, 4, 256):
            _lib.call("vkh_reduce_set_config", gpu._handle, 1, 1, B, d, -1)


def test_graph_precision_per_product_family(gpu):
    """precision given per product family: all three "bf16" is the same step
    as precision="bf16"; a mix (split-BF16 weight gradients) captured equals
    the same kernels launched one at a time, and trains."""
    dims, acts = [(32, 128), (128, 128), (128, 1)], ["relu", "relu", None]
    rng = np.random.default_rng(17)
    x = rng.standard_normal((256, 32)).astype(np.float32)
    y = rng.standard_normal((256, 1)).astype(np.float32)
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    same = {"forward": "bf16", "input_grad": "bf16", "weight_grad": "bf16"}
    mixed = {"forward": "bf16", "input_grad": "bf16", "weight_grad": "bf16x3"}
    nets = [nn.GraphSequence(make_net(gpu, dims, acts, 40, lambda: nn.Ridge(1e-4)), nn.HuberLoss(),
                             precision=p, capture=c)
            for p, c in (("bf16", True), (same, True), (mixed, True), (mixed, False))]
    losses = []
    for _ in range(4):
        ls = [np.asarray(n.train(X, Y)[1]).item() for n in nets]
        assert ls[0] == ls[1] and ls[2] == ls[3], ls
        losses.append(ls[2])
    for a, b in zip(weights(nets[0].L), weights(nets[1].L)):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(weights(nets[2].L), weights(nets[3].L)):
        np.testing.assert_array_equal(a, b)
    with pytest.raises(ValueError):
        nn.GraphSequence(make_net(gpu, dims, acts, 40), nn.HuberLoss(), precision={"backward": "bf16"})
