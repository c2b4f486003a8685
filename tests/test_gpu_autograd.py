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
