"""torch.autograd wrappers (vulkpy_amd.torch_autograd) on the GPU.

* FP64 central finite differences of the forward functions, written out in
  float64 numpy, against the analytic gradients from the HIP backward kernels
  on a reduced copy of the C4 network (batch 8, widths 16);
* float64 torch autograd on the CPU as a second reference, at larger sizes;
* Dense against vulkpy's own Dense.backward (nn.Dense on vk.Array): the same
  kernels, so the gradients are identical bit for bit.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import vulkpy_amd as vk  # noqa: E402
from vulkpy_amd import nn  # noqa: E402
from vulkpy_amd import torch_autograd as ta  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("torch sees no ROCm GPU")
    return torch.device("cuda", 0)


# ---- float64 forward of the reduced C4 network ----------------------------------------
def f64_softmax(x):
    e = np.exp(x - x.max(axis=1, keepdims=True))
    return e / e.sum(axis=1, keepdims=True)


def f64_net_loss(params, x, t):
    W1, b1, W2, b2, W3, b3 = params
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = 1.0 / (1.0 + np.exp(-(h1 @ W2.T + b2)))
    z = h2 @ W3.T + b3
    return np.mean(np.sum(-t * np.log(f64_softmax(z) + 1e-8), axis=1))


def fd_grad(f, params, k, h=1e-6):
    p = [q.copy() for q in params]
    g = np.zeros_like(p[k])
    it = np.nditer(p[k], flags=["multi_index"])
    for _ in it:
        i = it.multi_index
        old = p[k][i]
        p[k][i] = old + h
        fp = f(p)
        p[k][i] = old - h
        fm = f(p)
        p[k][i] = old
        g[i] = (fp - fm) / (2 * h)
    return g


def test_fd_reduced_c4_network(dev):
    rng = np.random.default_rng(21)
    B, D, H, O = 8, 16, 16, 10
    x = rng.standard_normal((B, D)).astype(np.float32)
    t = np.identity(O, dtype=np.float32)[rng.integers(0, O, B)]
    shapes = [(H, D), (H,), (H, H), (H,), (O, H), (O,)]
    params = [(rng.standard_normal(s) * (0.5 if len(s) == 2 else 0.1)).astype(np.float32) for s in shapes]

    tp = [torch.tensor(p, device=dev, requires_grad=True) for p in params]
    X, T = torch.tensor(x, device=dev), torch.tensor(t, device=dev)
    h1 = ta.relu(ta.dense(X, tp[0], tp[1]))
    h2 = ta.sigmoid(ta.dense(h1, tp[2], tp[3]))
    L = ta.softmax_cross_entropy(ta.dense(h2, tp[4], tp[5]), T)
    L.backward()
    p64 = [p.astype(np.float64) for p in params]
    x64, t64 = x.astype(np.float64), t.astype(np.float64)
    assert abs(L.item() - f64_net_loss(p64, x64, t64)) <= 1e-5 * abs(L.item()) + 1e-6
    for k in range(6):
        fd = fd_grad(lambda p: f64_net_loss(p, x64, t64), p64, k)
        got = tp[k].grad.cpu().numpy().astype(np.float64)
        np.testing.assert_allclose(got, fd, rtol=1e-3, atol=1e-5 * np.abs(fd).max() + 1e-7,
                                   err_msg=f"parameter {k}")


@pytest.mark.parametrize("name", ["mse", "huber", "softmax"])
def test_fd_losses_and_softmax(dev, name):
    rng = np.random.default_rng(5)
    B, C = 8, 6
    x = (rng.standard_normal((B, C)) * 1.5).astype(np.float32)
    t = rng.standard_normal((B, C)).astype(np.float32)
    w = rng.standard_normal((B, C)).astype(np.float32)  # projection for softmax
    X = torch.tensor(x, device=dev, requires_grad=True)
    T = torch.tensor(t, device=dev)
    if name == "mse":
        L = ta.mse(X, T)
        f = lambda v: np.mean(np.sum((t.astype(np.float64) - v) ** 2, axis=1))  # noqa: E731
    elif name == "huber":
        L = ta.huber(X, T, reduce="sum")
        f = lambda v: np.sum(0.5 * np.minimum((t - v) ** 2, np.abs(t - v)))  # noqa: E731
    else:
        L = (ta.softmax(X) * torch.tensor(w, device=dev)).sum()
        f = lambda v: np.sum(f64_softmax(v) * w.astype(np.float64))  # noqa: E731
    L.backward()
    fd = fd_grad(lambda p: f(p[0]), [x.astype(np.float64)], 0)
    np.testing.assert_allclose(X.grad.cpu().numpy(), fd, rtol=1e-4, atol=1e-6)


def test_against_torch_float64(dev):
    """Forward and all gradients against float64 torch autograd on the CPU."""
    rng = np.random.default_rng(9)
    B, D, H, O = 300, 70, 129, 7
    x = rng.standard_normal((B, D)).astype(np.float32)
    t = np.identity(O, dtype=np.float32)[rng.integers(0, O, B)]
    W1 = (rng.standard_normal((H, D)) * 0.2).astype(np.float32)
    b1 = (rng.standard_normal(H) * 0.1).astype(np.float32)
    W2 = (rng.standard_normal((O, H)) * 0.2).astype(np.float32)
    b2 = (rng.standard_normal(O) * 0.1).astype(np.float32)

    def run(mod, device, dtype):
        ps = [torch.tensor(a, device=device, dtype=dtype, requires_grad=True) for a in (x, W1, b1, W2, b2)]
        X, w1, c1, w2, c2 = ps
        T = torch.tensor(t, device=device, dtype=dtype)
        if mod is ta:
            z = ta.dense(ta.sigmoid(ta.dense(X, w1, c1)), w2, c2)
            L = ta.softmax_cross_entropy(z, T) + ta.mse(ta.softmax(z), T, reduce="sum")
        else:
            z = torch.nn.functional.linear(torch.sigmoid(torch.nn.functional.linear(X, w1, c1)), w2, c2)
            sm = torch.softmax(z, dim=1)
            L = (-(T * torch.log(sm + 1e-8)).sum(1)).mean() + ((T - sm) ** 2).sum()
        L.backward()
        return [L.detach()] + [p.grad for p in ps]

    got = run(ta, dev, torch.float32)
    want = run(torch, "cpu", torch.float64)
    for k, (g, w) in enumerate(zip(got, want)):
        g = g.cpu().double().numpy()
        w = w.numpy()
        np.testing.assert_allclose(g, w, rtol=2e-4, atol=2e-5 * max(np.abs(w).max(), 1e-3),
                                   err_msg=f"output {k}")


@pytest.mark.parametrize("precision,tol", [("bf16", 1e-2), ("bf16x3", 1e-5)])
def test_dense_precision_modes(dev, precision, tol):
    rng = np.random.default_rng(3)
    B, D, O = 512, 384, 256
    x = rng.standard_normal((B, D)).astype(np.float32)
    w = rng.standard_normal((O, D)).astype(np.float32)
    dy = rng.standard_normal((B, O)).astype(np.float32)
    X = torch.tensor(x, device=dev, requires_grad=True)
    Wt = torch.tensor(w, device=dev, requires_grad=True)
    y = ta.dense(X, Wt, None, precision=precision)
    y.backward(torch.tensor(dy, device=dev))
    x64, w64, dy64 = x.astype(np.float64), w.astype(np.float64), dy.astype(np.float64)
    for got, want in ((y.detach(), x64 @ w64.T), (X.grad, dy64 @ w64), (Wt.grad, dy64.T @ x64)):
        got = got.cpu().double().numpy()
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err <= tol, (precision, err)


def test_dense_matches_vulkpy_backward(dev, gpu):
    """The wrapper's gradients equal nn.Dense.backward's (same kernels)."""
    rng = np.random.default_rng(13)
    B, D, O = 96, 40, 24
    x = rng.standard_normal((B, D)).astype(np.float32)
    dy = rng.standard_normal((B, O)).astype(np.float32)
    layer = nn.Dense(gpu, D, O, w_init=nn.HeNormal(gpu, D, seed=1))
    layer.b.value[:] = rng.standard_normal(O).astype(np.float32)
    y_vk = layer(vk.Array(gpu, data=x))
    layer.zero_grad()
    dx_vk = layer.backward(vk.Array(gpu, data=dy))
    gpu.wait()
    W = torch.tensor(np.asarray(layer.w.value), device=dev, requires_grad=True)
    b = torch.tensor(np.asarray(layer.b.value), device=dev, requires_grad=True)
    X = torch.tensor(x, device=dev, requires_grad=True)
    y = ta.dense(X, W, b)
    y.backward(torch.tensor(dy, device=dev))
    np.testing.assert_array_equal(y.detach().cpu().numpy(), np.asarray(y_vk))
    np.testing.assert_array_equal(X.grad.cpu().numpy(), np.asarray(dx_vk))
    np.testing.assert_array_equal(W.grad.cpu().numpy(), np.asarray(layer.w.grad))
    np.testing.assert_array_equal(b.grad.cpu().numpy(), np.asarray(layer.b.grad))


def f64_relu(x):
    return np.maximum(x, 0.0)


def test_c4_full_size_autograd_matches_vulkpy_backward(dev, gpu):
    """The config-4 network at full size (784 -> 1024 -> 1024 -> 10, ReLU,
    Sigmoid, softmax cross entropy, batch 4096): the torch.autograd wrappers'
    gradients equal those of vulkpy's own layer-by-layer backward (nn.Dense /
    ReLU / Sigmoid / SoftmaxCrossEntropyLoss on vk.Array) bit for bit.  The
    forward is within 1e-4 of a float64 forward of the same weights, and the
    gradients within 1e-4 (relative to each gradient's largest entry) of a
    float64 backward from the GPU's own activations (a float64 forward would
    flip ReLU masks where a pre-activation is within FP32 rounding of 0)."""
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
    h1_vk = relu(layers[0](X))
    h2_vk = sig(layers[1](h1_vk))
    z = layers[2](h2_vk)
    loss_vk = float(np.asarray(lossf(z, T)))
    h1g, h2g, zg = (np.asarray(a).astype(np.float64) for a in (h1_vk, h2_vk, z))
    dz = lossf.grad()
    layers[0].backward(relu.backward(layers[1].backward(sig.backward(layers[2].backward(dz)))))
    gpu.wait()
    vk_grads = [np.asarray(p.grad).copy() for L in layers for p in (L.w, L.b)]
    weights = [np.asarray(p.value).copy() for L in layers for p in (L.w, L.b)]

    P = [torch.tensor(w, device=dev, requires_grad=True) for w in weights]
    Xt, Tt = torch.tensor(x, device=dev), torch.tensor(t, device=dev)
    zt = ta.dense(ta.sigmoid(ta.dense(ta.relu(ta.dense(Xt, P[0], P[1])), P[2], P[3])), P[4], P[5])
    loss_t = ta.softmax_cross_entropy(zt, Tt)
    loss_t.backward()
    torch.cuda.synchronize()
    for k, (p, g) in enumerate(zip(P, vk_grads)):
        np.testing.assert_array_equal(p.grad.cpu().numpy(), g, err_msg=f"parameter {k}")
    assert abs(float(loss_t.item()) - loss_vk) <= 1e-6 * abs(loss_vk)

    # float64 forward of the same weights
    W1, b1, W2, b2, W3, b3 = [w.astype(np.float64) for w in weights]
    x64, t64 = x.astype(np.float64), t.astype(np.float64)
    h1 = f64_relu(x64 @ W1.T + b1)
    h2 = 1.0 / (1.0 + np.exp(-(h1 @ W2.T + b2)))
    z64 = h2 @ W3.T + b3
    for k, (g, r) in enumerate(((h1g, h1), (h2g, h2), (zg, z64))):
        assert np.abs(g - r).max() <= 1e-4 * np.abs(r).max(), ("forward", k)
    sm = f64_softmax(z64)
    loss64 = np.mean(np.sum(-t64 * np.log(sm + 1e-8), axis=1))
    assert abs(loss_vk - loss64) <= 1e-5 * abs(loss64)
    # float64 backward from the GPU's activations
    d3 = (f64_softmax(zg) - t64) / B
    d2 = (d3 @ W3) * h2g * (1.0 - h2g)
    d1 = (d2 @ W2) * (h1g > 0)
    ref = [d1.T @ x64, d1.sum(0), d2.T @ h1g, d2.sum(0), d3.T @ h2g, d3.sum(0)]
    for k, (g, r) in enumerate(zip(vk_grads, ref)):
        assert np.abs(g - r).max() <= 1e-4 * np.abs(r).max(), ("gradient", k)


def test_rejects_bad_inputs(dev):
    x = torch.zeros((4, 3), device=dev, dtype=torch.float64)
    with pytest.raises(TypeError):
        ta.relu(x)
    with pytest.raises(ValueError):
        ta.dense(torch.zeros((4, 3), device=dev), torch.zeros((2, 5), device=dev))
    with pytest.raises(ValueError):
        ta.mse(torch.zeros((4, 3), device=dev), torch.zeros((4, 2), device=dev))
    with pytest.raises(ValueError):
        ta.dense(torch.zeros((4, 3), device=dev), torch.zeros((2, 3), device=dev), precision="fp16")
