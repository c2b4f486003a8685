"""The nn workloads at their full sizes.

* The regression MLP 256 -> 4096 -> 4096 -> 1 (ReLU, HuberLoss, Ridge(1e-4)
  on the weights, Adam lr=1e-3, batch 16384), three steps:
  - the FP32 step captured as one hipGraph (nn.GraphSequence) equals the
    eager Sequence.train(regularize=True) step bit for bit (prediction, loss,
    every parameter), and each stage of the captured step is checked against
    the oracle's restatement of the reference's formulas, from the stage's
    own inputs: forward products and the input-gradient product on sampled
    rows and the weight gradients on sampled output units against FP64 within
    the length-k FP32 bound; Huber rows (nn/losses.py:365-369), the Huber
    gradient (391-393, scaled by 1/B) and the ReLU derivative of the
    K = 1 product exactly; the batch mean within its bound; the Adam update
    (nn/optimizers.py:235-253) of every parameter bit for bit;
  - the BF16 step captured as one hipGraph equals the eager Sequence whose
    Dense layers run in BF16, bit for bit.
* The classifier 784 -> 1024 -> 1024 -> 10 (ReLU, Sigmoid, softmax cross
  entropy, batch 4096): the layer-by-layer backward's weight and bias
  gradients against the reference's own broadcast-product formula
  (nn/layers.py:126-141, restated by oracle.dense_backward_reference,
  sequential over the batch in FP32) on sampled output units, from the
  gradients and activations each layer saw.
"""
import gc

import numpy as np
import pytest

import vulkpy_amd as vk
from vulkpy_amd import nn
from oracle import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
f32 = np.float32
U = 2.0 ** -24


def assert_bits(got, want, what=""):
    got = np.asarray(got, dtype=f32)
    want = np.asarray(want, dtype=f32)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    if not same.all():
        bad = np.argwhere(~same)
        raise AssertionError(f"{what}: {bad.shape[0]} mismatches, first at {bad[:3].tolist()}")


def within(got, exact, bound, what):
    err = np.abs(np.asarray(got, dtype=np.float64) - exact)
    ok = err <= bound
    assert ok.all(), f"{what}: {int((~ok).sum())} outside the bound, worst excess {float((err - bound).max())}"


B5, D5, H5 = 16384, 256, 4096


def c5_data():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((B5, D5), dtype=np.float32)
    proj = (rng.standard_normal((D5, 1)) / np.sqrt(D5)).astype(np.float32)
    y = (np.sin(x @ proj) + 0.1 * rng.standard_normal((B5, 1))).astype(np.float32)
    return x, y


def c5_layers(gpu, precision):
    out = []
    dims = [(D5, H5), (H5, H5), (H5, 1)]
    for k, (i, o) in enumerate(dims):
        out.append(nn.Dense(gpu, i, o, w_init=nn.HeNormal(gpu, i, seed=100 + k),
                            w_opt=nn.Adam(gpu, lr=1e-3), b_opt=nn.Adam(gpu, lr=1e-3),
                            w_reg=nn.Ridge(1e-4), precision=precision))
        if k + 1 < len(dims):
            out.append(nn.ReLU())
    return out


def dense_of(layers):
    return [L for L in layers if isinstance(L, nn.Dense)]


def params_of(layers):
    return [p for L in dense_of(layers) for p in (L.w, L.b)]


def host(a):
    return np.array(np.asarray(a), dtype=f32, copy=True)


def test_c5_captured_fp32_step_vs_eager_and_oracle(gpu):
    x, y = c5_data()
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    eager_layers, graph_layers = c5_layers(gpu, "fp32"), c5_layers(gpu, "fp32")
    eager = nn.Sequence(eager_layers, nn.HuberLoss(), regularize=True)
    graph = nn.GraphSequence(graph_layers, nn.HuberLoss(), precision="fp32", autotune=True)
    rng = np.random.default_rng(0)
    rows = np.sort(rng.choice(B5, 48, replace=False))
    units = np.sort(rng.choice(H5, 24, replace=False))
    c2 = float(f32(2 * 1e-4))                          # Ridge gradient factor, rounded as the kernel takes it
    scale = float(f32(1 / B5))
    for step in range(3):
        ps = params_of(graph_layers)
        before = [(host(p.value), host(p.opt_state.m), host(p.opt_state.v)) for p in ps]
        pe, le = eager.train(X, Y)
        pg, lg = graph.train(X, Y)
        gpu.wait()
        # captured == eager
        assert np.asarray(le).item() == np.asarray(lg).item(), step
        assert_bits(pg, pe, f"prediction {step}")
        for k, (a, b) in enumerate(zip(params_of(eager_layers), ps)):
            assert_bits(b.value, a.value, f"parameter {k} step {step}")

        plan = graph._plan
        (W1, b1), (W2, b2), (W3, b3) = [(before[2 * i][0], before[2 * i + 1][0]) for i in range(3)]
        h1, h2, out = (host(a) for a in plan.hs)
        d1, d2, d3 = (host(a) for a in plan.ds)
        # forward, each product from the GPU's own input (dense.forward, layers.py:71-81)
        x64, h1r, h2r = x[rows].astype(np.float64), h1[rows].astype(np.float64), h2[rows].astype(np.float64)
        for name, inp, W, b, got, K, act in (("h1", x64, W1, b1, h1[rows], D5, True),
                                             ("h2", h1r, W2, b2, h2[rows], H5, True),
                                             ("out", h2r, W3, b3, out[rows], H5, False)):
            pre = inp @ W.astype(np.float64).T + b
            bound = (K + 2) * U * (np.abs(inp) @ np.abs(W.astype(np.float64)).T + np.abs(b))
            within(got, np.maximum(pre, 0) if act else pre, bound, f"{name} step {step}")
        # loss rows, batch mean and the loss gradient
        L = O.huber_rows(out, y)
        assert_bits(plan.L_rows, L, "Huber rows")
        m64 = L.astype(np.float64).mean()
        assert abs(np.asarray(lg).item() - m64) <= (B5 + 2) * U * m64, step
        assert_bits(d3, O.loss_grad("huber", out, y, scale=scale), "Huber gradient")
        # input gradients: the K = 1 product dY3 W3 with ReLU' is exact; the
        # 4096-deep one within its bound
        want2 = O.relu_backward(h2, (d3 * W3).astype(f32))
        assert np.array_equal(d2, want2), "dY2"
        g = d2[rows].astype(np.float64) @ W2.astype(np.float64)
        bound = (H5 + 1) * U * (np.abs(d2[rows]).astype(np.float64) @ np.abs(W2).astype(np.float64))
        within(d1[rows], g * (h1[rows] > 0), bound, f"dY1 step {step}")
        # parameter gradients on sampled output units: grad = 2c w + dY^T x
        grads = [host(p.grad) for p in ps]
        for i, (dy, inp, W, sel) in enumerate(((d1, x, W1, units), (d2, h1, W2, units),
                                               (d3, h2, W3, np.array([0])))):
            dys, inp64 = dy[:, sel].astype(np.float64), inp.astype(np.float64)
            exact = c2 * W[sel].astype(np.float64) + dys.T @ inp64
            bound = (B5 + 2) * U * (np.abs(dys).T @ np.abs(inp64)) + 2 * U * np.abs(c2 * W[sel])
            within(grads[2 * i][sel], exact, bound, f"dW{i + 1} step {step}")
            db = dy.astype(np.float64).sum(axis=0)
            within(grads[2 * i + 1], db, (B5 + 1) * U * np.abs(dy).astype(np.float64).sum(axis=0),
                   f"db{i + 1} step {step}")
            del dys, inp64
        # Adam, every element of every parameter
        for k, p in enumerate(ps):
            st = p.opt_state
            o = st.opt
            w0, m0, v0 = before[k]
            w1, m1, v1 = O.adam(w0, grads[k], m0, v0, o.lr, o.beta1, o.beta2, o.eps, st.beta1t, st.beta2t)
            assert_bits(p.value, w1, f"Adam w {k} step {step}")
            assert_bits(st.m, m1, f"Adam m {k} step {step}")
            assert_bits(st.v, v1, f"Adam v {k} step {step}")
        del h1, h2, d1, d2, grads, before
        gc.collect()


def test_c5_captured_bf16_step_equals_eager_bf16(gpu):
    x, y = c5_data()
    X, Y = vk.Array(gpu, data=x), vk.Array(gpu, data=y)
    eager_layers, graph_layers = c5_layers(gpu, "bf16"), c5_layers(gpu, "bf16")
    eager = nn.Sequence(eager_layers, nn.HuberLoss(), regularize=True)
    graph = nn.GraphSequence(graph_layers, nn.HuberLoss(), precision="bf16", autotune=True)
    for step in range(3):
        pe, le = eager.train(X, Y)
        pg, lg = graph.train(X, Y)
        gpu.wait()
        assert np.asarray(le).item() == np.asarray(lg).item(), step
        assert_bits(pg, pe, f"prediction {step}")
        for k, (a, b) in enumerate(zip(params_of(eager_layers), params_of(graph_layers))):
            assert_bits(b.value, a.value, f"parameter {k} step {step}")


def test_c4_backward_against_reference_formula(gpu):
    """Each Dense.backward of the full-size classifier against the reference's
    formula db = sum_b dy, dW = sum_b dy[b, :, None] x[b, None, :] evaluated
    sequentially over the batch in FP32 (oracle.dense_backward_reference),
    on 32 sampled output units of each layer, from the dy and x that layer
    saw.  The two sum the batch in different orders, so they agree within
    twice the length-B FP32 bound; dx against FP64."""
    rng = np.random.default_rng(44)
    B, dims, C = 4096, [(784, 1024), (1024, 1024), (1024, 10)], 10
    x = rng.standard_normal((B, 784)).astype(np.float32)
    t = np.identity(C, dtype=np.float32)[rng.integers(0, C, B)]
    layers = [nn.Dense(gpu, i, o, w_init=nn.HeNormal(gpu, i, seed=70 + k)) for k, (i, o) in enumerate(dims)]
    for k, L in enumerate(layers):
        L.b.value[:] = (0.1 * rng.standard_normal(dims[k][1])).astype(np.float32)
        L.zero_grad()
    relu, sig, lossf = nn.ReLU(), nn.Sigmoid(), nn.SoftmaxCrossEntropyLoss()
    X, T = vk.Array(gpu, data=x), vk.Array(gpu, data=t)
    h1 = relu(layers[0](X))
    h2 = sig(layers[1](h1))
    z = layers[2](h2)
    lossf(z, T)
    dy3 = lossf.grad()
    dx3 = layers[2].backward(dy3)
    dy2 = sig.backward(dx3)
    dx2 = layers[1].backward(dy2)
    dy1 = relu.backward(dx2)
    layers[0].backward(dy1)
    gpu.wait()
    inputs = [x, host(h1), host(h2)]
    dys = [host(dy1), host(dy2), host(dy3)]
    dxs = [None, host(dx2), host(dx3)]
    for k, L in enumerate(layers):
        W = host(L.w.value)
        gW, gb = host(L.w.grad), host(L.b.grad)
        sel = np.sort(rng.choice(dims[k][1], min(32, dims[k][1]), replace=False))
        dW_ref, db_ref, _ = O.dense_backward_reference(inputs[k], dys[k][:, sel], W[sel])
        a64 = np.abs(dys[k][:, sel]).astype(np.float64).T @ np.abs(inputs[k]).astype(np.float64)
        within(gW[sel], dW_ref.astype(np.float64), 2 * (B + 1) * U * a64 + 1e-45, f"dW layer {k + 1}")
        within(gb[sel], db_ref.astype(np.float64),
               2 * (B + 1) * U * np.abs(dys[k][:, sel]).astype(np.float64).sum(axis=0) + 1e-45,
               f"db layer {k + 1}")
        if dxs[k] is not None:
            dx64 = dys[k].astype(np.float64) @ W.astype(np.float64)
            bound = (dims[k][1] + 1) * U * (np.abs(dys[k]).astype(np.float64) @ np.abs(W).astype(np.float64))
            within(dxs[k], dx64, bound, f"dx layer {k + 1}")
