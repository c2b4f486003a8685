"""GPU parity of the nn path: the reference's test/test_nn.py known answers,
fused kernels against the oracle's NumPy restatement of the reference's FP32
operation order (bit for bit where the operations are +, -, *, /, sqrt), and a
reduced training run against vulkpy's own backward formulas."""
import numpy as np
import pytest

import vulkpy_amd as vk
from vulkpy_amd import nn
from vulkpy_amd.nn.parameters import Parameter
from oracle import oracle as O

pytestmark = pytest.mark.gpu
f32 = np.float32


def A(gpu, x):
    return vk.Array(gpu, data=np.asarray(x, dtype=np.float32))


def bits_equal(got, want):
    got = np.asarray(got, dtype=f32)
    want = np.asarray(want, dtype=f32)
    assert got.shape == want.shape, (got.shape, want.shape)
    bad = ~((got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want)))
    assert not bad.any(), f"{bad.sum()} mismatches; max diff {np.nanmax(np.abs(got - want))}"


def within_ulp(got, want, ulps=1):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=f32)
    tol = np.spacing(np.abs(want)).astype(np.float64) * ulps
    assert np.all(np.abs(got - want) <= tol + 1e-45), np.abs(got - want).max()


# ---- test/test_nn.py known answers ---------------------------------------------------
def test_initializers(gpu):
    np.testing.assert_allclose(nn.Constant(0.0)(gpu, (3, 1)), [[0.0], [0.0], [0.0]])
    he = nn.HeNormal(gpu, input_dim=2, seed=645)
    assert he.stddev == 1.0
    rng = vk.random.Xoshiro128pp(gpu, seed=645)
    np.testing.assert_allclose(he(gpu, (10,)), rng.normal(shape=(10,)))


def test_optimizers(gpu):
    sgd = nn.SGD(lr=0.01)
    grad = A(gpu, [1, 2, 3])
    np.testing.assert_allclose(sgd.init_state(grad.shape).grad2diff(grad), grad * (-0.01))
    adam = nn.Adam(gpu)
    state = adam.init_state(grad.shape)
    assert state.beta1t == 1.0 and state.beta2t == 1.0
    state.grad2diff(grad)
    assert state.beta1t == adam.beta1 and state.beta2t == adam.beta2
    nn.AdaGrad(gpu).init_state(grad.shape).grad2diff(grad)


def test_layers(gpu):
    relu = nn.ReLU()
    x = A(gpu, [[-0.2, 0.0, 0.2]])
    np.testing.assert_allclose(relu(x), [[0.0, 0.0, 0.2]])
    np.testing.assert_allclose(relu.backward(A(gpu, [[0.7, 0.8, 0.9]])), [[0.0, 0.0, 0.9]])
    d = np.asarray([[-100, -0.1, 0, 10, 100]])
    sig = nn.Sigmoid()
    y = sig(A(gpu, d))
    np.testing.assert_allclose(y, 1 / (1 + np.exp(-d)), rtol=1e-7, atol=1e-7)
    dy = A(gpu, [[0.1, 0.2, 0.3, 0.5, 0.7]])
    np.testing.assert_allclose(sig.backward(dy), dy * y * (1 - y))
    sm = nn.Softmax()
    np.testing.assert_allclose(sm(A(gpu, [[1.0, 1.0]])), [[0.5, 0.5]])
    np.testing.assert_allclose(sm(A(gpu, [[100.0, 0.0]])), [[1.0, 0.0]])
    y = sm(A(gpu, d))
    e = np.exp(d - d.max(axis=1))
    np.testing.assert_allclose(y, e / e.sum(axis=1, keepdims=True), rtol=1e-7, atol=1e-7)
    np.testing.assert_allclose(sm.backward(dy), dy * y * (1 - y))
    with pytest.raises(ValueError):
        relu(A(gpu, [1, 2]))


def test_dense(gpu):
    dense = nn.Dense(gpu, 2, 2, w_init=nn.Constant(0.0))
    np.testing.assert_allclose(dense(A(gpu, [[1, 2], [3, 4]])), [[0, 0], [0, 0]])
    dense = nn.Dense(gpu, 2, 2, w_init=nn.Constant(0.0), b_init=nn.Constant(1.0))
    np.testing.assert_allclose(dense(A(gpu, [[1, 2], [3, 4]])), [[1, 1], [1, 1]])
    dense = nn.Dense(gpu, 2, 2)
    y = dense(A(gpu, [[2, 3], [2, 3]]))
    np.testing.assert_allclose(y[0, :], y[1, :])
    dense = nn.Dense(gpu, 2, 2)
    np.testing.assert_allclose(dense.w.grad, [[0, 0], [0, 0]])
    np.testing.assert_allclose(dense.b.grad, [0, 0])
    dense(A(gpu, [[1, 2], [3, 4]]))
    dy = A(gpu, [[4, 2], [1, 3]])
    dx = dense.backward(dy)
    np.testing.assert_allclose(dense.w.grad, [[7, 12], [11, 16]])
    np.testing.assert_allclose(dense.b.grad, [5, 5])
    _w = np.asarray(dense.w.value)
    np.testing.assert_allclose(dx, np.asarray(dy) @ _w, rtol=1e-6)


def test_losses(gpu):
    L = nn.CrossEntropyLoss()
    np.testing.assert_allclose(L(A(gpu, [[1.0, 0.0]]), A(gpu, [[1.0, 0.0]])), [0.0])
    np.testing.assert_allclose(L(A(gpu, [[0.5, 0.5]]), A(gpu, [[0.5, 0.5]])), [0.6931472])
    _x = np.asarray([[0.7, 0.3], [0.2, 0.8], [1.0, 0.0]])
    _y = np.asarray([[1.0, 0.0], [0.0, 1.0], [1.0, 0.0]])
    _L = np.sum(-_y * np.log(_x + 1e-8), axis=1)
    _dx = -_y / (_x + 1e-8)
    for red, Lr, scale in (("mean", _L.mean(), 1 / 3), ("sum", _L.sum(), 1.0)):
        loss = nn.CrossEntropyLoss(reduce=red)
        np.testing.assert_allclose(loss(A(gpu, _x), A(gpu, _y)), Lr, atol=1e-7, rtol=1e-7)
        np.testing.assert_allclose(loss.grad(), _dx * scale, rtol=1e-6)
    sce = nn.SoftmaxCrossEntropyLoss()
    np.testing.assert_allclose(sce(A(gpu, [[100.0, 0.0]]), A(gpu, [[1.0, 0.0]])), [0.0])
    _x = np.asarray([[-1, 0], [10, 15]])
    _y = np.asarray([[1, 0], [0, 1]])
    e = np.exp(_x - _x.max(axis=1, keepdims=True))
    sm = e / e.sum(axis=1, keepdims=True)
    _L = (-_y * np.log(sm)).sum(axis=1)
    for red, Lr, scale in (("mean", _L.mean(axis=0), 0.5), ("sum", _L.sum(axis=0), 1.0)):
        sce = nn.SoftmaxCrossEntropyLoss(reduce=red)
        np.testing.assert_allclose(sce(A(gpu, _x), A(gpu, _y)), Lr, atol=1e-7, rtol=1e-7)
        np.testing.assert_allclose(sce.grad(), (sm - _y) * scale, atol=1e-7, rtol=1e-7)
    _x = np.asarray([[4, 2], [1, 1.5]])
    _y = np.asarray([[3, 2.2], [0.7, 1.5]])
    for red, red_f, gscale in (("mean", np.mean, 1.0), ("sum", np.sum, 2.0)):
        mse = nn.MSELoss(reduce=red)
        np.testing.assert_allclose(mse(A(gpu, _x), A(gpu, _y)), red_f(np.square(_y - _x).sum(axis=1)),
                                   atol=1e-7, rtol=1e-7)
        np.testing.assert_allclose(mse.grad(), gscale * (_x - _y), atol=1e-7, rtol=1e-7)
    _x = np.asarray([[1.0, 2.2], [-3.0, 0.7]])
    _y = np.asarray([[10, 3.0], [-5, 0.5]])
    for red, Lr, g in (("mean", 2.92, [[-0.5, -0.4], [0.5, 0.1]]), ("sum", 5.84, [[-1.0, -0.8], [1.0, 0.2]])):
        hub = nn.HuberLoss(reduce=red)
        np.testing.assert_allclose(hub(A(gpu, _x), A(gpu, _y)), [Lr])
        np.testing.assert_allclose(hub.grad(), g)
    mix = nn.MixLoss([(0.5, nn.MSELoss()), (2.0, nn.HuberLoss())])
    Lm = mix(A(gpu, _x), A(gpu, _y))
    np.testing.assert_allclose(Lm, 0.5 * np.square(_y - _x).sum(axis=1).mean() + 2 * 2.92, rtol=1e-6)
    with pytest.raises(ValueError):
        nn.MixLoss([])


def test_regularizers(gpu):
    for v in (0.0, 3.5, -3.5):
        p = Parameter(gpu, (1,), initializer=nn.Constant(v))
        np.testing.assert_allclose(nn.Ridge(1.0).loss(p.value), np.asarray(v ** 2))
        np.testing.assert_allclose(nn.Ridge(1.0).grad(p.value), np.asarray((2 * v,)))
        np.testing.assert_allclose(nn.Lasso(1.0).loss(p.value), np.asarray(abs(v)))
        np.testing.assert_allclose(nn.Lasso(1.0).grad(p.value), np.asarray((np.sign(v),)))
        np.testing.assert_allclose(nn.Elastic(1.0, 1.0).loss(p.value), np.asarray(v ** 2 + abs(v)))
        np.testing.assert_allclose(nn.Elastic(1.0, 1.0).grad(p.value), np.asarray((2 * v + np.sign(v),)))


@pytest.mark.parametrize("kind", ["lasso", "ridge", "elastic"])
def test_regularizer_grad_fused(gpu, kind):
    """Regularizer.grad is one kernel, bit for bit the oracle's FP32 restatement
    of the reference's composition; Parameter.regular_grad adds it onto the
    gradient in place, equal to grad + R.grad(w)."""
    rng = np.random.default_rng(12)
    w = rng.standard_normal(100003).astype(f32)
    w[:4] = [0.0, -0.0, 1e-41, np.inf]
    g0 = rng.standard_normal(w.shape).astype(f32)
    c1, c2 = 3e-3, 5e-4
    R = {"lasso": nn.Lasso(c1), "ridge": nn.Ridge(c2), "elastic": nn.Elastic(c1, c2)}[kind]
    want = O.regularizer_grad(kind, w, l1=c1, l2=c2)
    bits_equal(R.grad(A(gpu, w)), want)
    p = Parameter(gpu, w.shape, initializer=nn.Constant(0.0), regularizer=R)
    p.value[:] = w
    p.grad[:] = g0
    p.regular_grad()
    bits_equal(p.grad, (g0 + want).astype(f32))
    # composition with array ops, as the reference writes it
    W = A(gpu, w)
    comp = {"lasso": lambda: c1 * W.sign(), "ridge": lambda: (2 * c2) * W,
            "elastic": lambda: c1 * W.sign() + (2 * c2) * W}[kind]()
    bits_equal(R.grad(W), comp)


# ---- fused kernels against the oracle -------------------------------------------------------
def test_activation_kernels_bit_exact(gpu):
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((257, 33)) * 4).astype(f32)
    x[0, :5] = [0.0, -0.0, 1e-40, -1e-40, 88.0]
    dy = rng.standard_normal(x.shape).astype(f32)
    relu = nn.ReLU()
    bits_equal(relu(A(gpu, x)), O.relu(x))
    bits_equal(relu.backward(A(gpu, dy)), O.relu_backward(O.relu(x), dy))
    sig = nn.Sigmoid()
    within_ulp(sig(A(gpu, x)), O.sigmoid(x), 1)
    y = np.asarray(sig._y)
    bits_equal(sig.backward(A(gpu, dy)), O.sigmoid_backward(y, dy))


def test_softmax_and_losses_vs_oracle(gpu):
    rng = np.random.default_rng(1)
    for cols in (10, 64, 65, 1000):
        x = (rng.standard_normal((300, cols)) * 3).astype(f32)
        lab = rng.integers(0, cols, 300)
        y = np.identity(cols, dtype=f32)[lab]
        sm = nn.Softmax()
        got = np.asarray(sm(A(gpu, x)))
        want = O.softmax(x)
        if cols <= 64:
            within_ulp(got, want, 1)  # same sequential order
        else:
            np.testing.assert_allclose(got, want, rtol=cols * 2 ** -23, atol=1e-30)
        sce = nn.SoftmaxCrossEntropyLoss()
        L = float(np.asarray(sce(A(gpu, x), A(gpu, y))))
        L_ref = float(O.batch_mean(O.cross_entropy_rows(O.softmax(x), y)))
        assert abs(L - L_ref) <= 1e-5 * max(1.0, abs(L_ref))
        g = np.asarray(sce.grad())
        np.testing.assert_allclose(g, O.loss_grad("sce", O.softmax(x), y, 1 / 300), atol=1e-7)
    x = rng.standard_normal((500, 3)).astype(f32)
    t = rng.standard_normal((500, 3)).astype(f32)
    mse, hub = nn.MSELoss(), nn.HuberLoss()
    bits_equal(mse.forward(A(gpu, x), A(gpu, t)), O.mse_rows(x, t))
    bits_equal(hub.forward(A(gpu, x), A(gpu, t)), O.huber_rows(x, t))
    mse(A(gpu, x), A(gpu, t))
    bits_equal(mse.grad(), O.loss_grad("mse", x, t, float(f32(1 / 500))))
    hub(A(gpu, x), A(gpu, t))
    bits_equal(hub.grad(), O.loss_grad("huber", x, t, float(f32(1 / 500))))


def test_optimizer_kernels_bit_exact(gpu):
    rng = np.random.default_rng(2)
    n = 10007
    w0 = rng.standard_normal(n).astype(f32)
    grads = [rng.standard_normal(n).astype(f32) for _ in range(4)]
    adam = nn.Adam(gpu, lr=1e-3)
    p = Parameter(gpu, (n,), opt=adam, initializer=lambda g, s: A(g, w0))
    w, m, v = w0, np.zeros(n, f32), np.zeros(n, f32)
    b1t = b2t = 1.0
    for g in grads:
        p.grad[:] = 0.0
        p.add_grad(A(gpu, g))
        p.update()
        b1t *= 0.9
        b2t *= 0.999
        w, m, v = O.adam(w, g, m, v, 1e-3, 0.9, 0.999, 1e-8, b1t, b2t)
    bits_equal(p.value, w)
    bits_equal(p.opt_state.m, m)
    bits_equal(p.opt_state.v, v)
    # the fused update equals the reference's composed grad2diff path
    q = Parameter(gpu, (n,), opt=nn.Adam(gpu, lr=1e-3), initializer=lambda g, s: A(g, w0))
    for g in grads:
        q.value += q.opt_state.grad2diff(A(gpu, g))
    bits_equal(q.value, w)
    sgd = Parameter(gpu, (n,), opt=nn.SGD(0.01), initializer=lambda g, s: A(g, w0))
    sgd.add_grad(A(gpu, grads[0]))
    sgd.update()
    bits_equal(sgd.value, O.sgd(w0, grads[0], 0.01))
    ada = Parameter(gpu, (n,), opt=nn.AdaGrad(gpu, lr=0.01, tau=0.5), initializer=lambda g, s: A(g, w0))
    ada.add_grad(A(gpu, grads[0]))
    ada.update()
    w2, h2 = O.adagrad(w0, grads[0], np.full(n, 0.5, f32), 0.01, 1e-8)
    bits_equal(ada.value, w2)
    bits_equal(ada.opt_state.h, h2)


def test_loss_rows_grad_equals_separate_kernels(gpu):
    """vkh_loss_rows_grad (one launch) against vkh_mse_rows / vkh_huber_rows
    and vkh_loss_grad, bit for bit, and against the oracle."""
    from vulkpy_amd import _lib
    h = gpu._handle
    rng = np.random.default_rng(4)
    for rows, cols in ((16384, 1), (513, 3)):
        x = (rng.standard_normal((rows, cols)) * 2).astype(f32)
        t = (rng.standard_normal((rows, cols)) * 2).astype(f32)
        X, T = A(gpu, x), A(gpu, t)
        scale = float(f32(1 / rows))
        for kind, rows_fn, name in ((1, "vkh_mse_rows", "mse"), (2, "vkh_huber_rows", "huber")):
            L1, G1, L2, G2 = (vk.Array(gpu, shape=s) for s in ((rows,), (rows, cols), (rows,), (rows, cols)))
            _lib.call("vkh_loss_rows_grad", h, kind, X._ptr, T._ptr, L1._ptr, G1._ptr, rows, cols, scale)
            _lib.call(rows_fn, h, X._ptr, T._ptr, L2._ptr, rows, cols)
            _lib.call("vkh_loss_grad", h, kind, X._ptr, T._ptr, G2._ptr, rows * cols, scale)
            gpu.wait()
            bits_equal(L1, np.asarray(L2))
            bits_equal(G1, np.asarray(G2))
            bits_equal(L1, O.mse_rows(x, t) if kind == 1 else O.huber_rows(x, t))
            bits_equal(G1, O.loss_grad(name, x, t, scale))


def test_adam_multi_equals_adam_device(gpu):
    """vkh_adam_device_multi on parameters of several sizes (aligned and
    ragged, with and without regularizer terms) against one vkh_adam_device
    each, bit for bit, over three steps with device beta powers."""
    from vulkpy_amd import _lib
    h = gpu._handle
    rng = np.random.default_rng(5)
    sizes = (1, 3, 4096, 10007, 262144 + 5, 17)
    regs = ((0, 0.0, 0.0), (2, 0.0, float(f32(2e-4))), (1, 1e-3, 0.0), (3, 1e-3, float(f32(2e-4))),
            (0, 0.0, 0.0), (2, 0.0, 0.5))
    def state():
        return [[A(gpu, rng.standard_normal(n)), A(gpu, np.zeros(n)), A(gpu, np.zeros(n))] for n in sizes]
    a, b = state(), state()
    for (wa, ma, va), (wb, mb, vb) in zip(a, b):
        wb[:] = wa
    # device beta powers: FP64 pairs held in uint32 buffers
    bpa = vk.U32Array(gpu, data=np.array([1.0, 1.0]).view(np.uint32))
    bpb = vk.U32Array(gpu, data=np.array([1.0, 1.0]).view(np.uint32))
    for step in range(3):
        grads = [A(gpu, rng.standard_normal(n)) for n in sizes]
        ga = [vk.Array(gpu, data=np.asarray(g)) for g in grads]
        gb = [vk.Array(gpu, data=np.asarray(g)) for g in grads]
        _lib.call("vkh_adam_pow_advance", h, bpa._ptr, 0.9, 0.999)
        _lib.call("vkh_adam_pow_advance", h, bpb._ptr, 0.9, 0.999)
        params = [_lib.AdamParam(w._ptr, g._ptr, m._ptr, v._ptr, n, 1e-3, 1e-8, 0.9, 0.999, bpa._ptr, *r)
                  for (w, m, v), g, n, r in zip(a, ga, sizes, regs)]
        _lib.call("vkh_adam_device_multi", h, (_lib.AdamParam * len(params))(*params), len(params))
        for (w, m, v), g, n, r in zip(b, gb, sizes, regs):
            _lib.call("vkh_adam_device", h, w._ptr, g._ptr, m._ptr, v._ptr, n, 1e-3, 0.9, 0.999, 1e-8,
                      bpb._ptr, None, *r)
        for arrs in (a, b, [[g] for g in ga], [[g] for g in gb]):
            for group in arrs:
                for x in group:
                    x._set_job([bpa, bpb])  # written through raw calls: the host mirrors are stale
        gpu.wait()
        for (wa, ma, va), (wb, mb, vb), g1, g2 in zip(a, b, ga, gb):
            for x, y in ((wa, wb), (ma, mb), (va, vb), (g1, g2)):
                bits_equal(x, np.asarray(y))


def test_dense_backward_vs_reference_formula(gpu):
    # vulkpy's Dense.backward (layers.py:126-141) restated by the oracle
    rng = np.random.default_rng(3)
    B, I, Out = 64, 48, 40
    x = rng.standard_normal((B, I)).astype(f32)
    dy = rng.standard_normal((B, Out)).astype(f32)
    d = nn.Dense(gpu, I, Out, w_init=nn.HeNormal(gpu, I, seed=1))
    d(A(gpu, x))
    dx = np.asarray(d.backward(A(gpu, dy)))
    W = np.asarray(d.w.value)
    dW, db, dx_ref = O.dense_backward_reference(x, dy, W)
    np.testing.assert_allclose(np.asarray(d.w.grad), dW, rtol=0, atol=B * 2 ** -20 * np.abs(dW).max())
    np.testing.assert_allclose(np.asarray(d.b.grad), db, rtol=0, atol=B * 2 ** -22 * np.abs(db).max())
    np.testing.assert_allclose(dx, dx_ref, rtol=0, atol=Out * 2 ** -20 * np.abs(dx_ref).max())
    y = np.asarray(d(A(gpu, x)))
    ref = O.batch_affine(W, np.asarray(d.b.value), x)
    np.testing.assert_allclose(y, ref, rtol=0, atol=I * 2 ** -22 * np.abs(ref).max())


def test_c4_reduced_training_matches_oracle(gpu):
    """Config 4 network at reduced width (B=64), 3 steps of Sequence.train.
    Each step: loss and gradients against a NumPy restatement of vulkpy's
    forward/backward (layers.py, losses.py) at the pre-step weights, then the
    Adam update bit for bit from the GPU's own gradients."""
    rng = np.random.default_rng(4)
    B, D0, D1, D2, C = 64, 20, 16, 12, 10
    x = rng.standard_normal((B, D0)).astype(f32)
    lab = rng.integers(0, C, B)
    y = np.identity(C, dtype=f32)[lab]
    dims = [(D0, D1), (D1, D2), (D2, C)]
    layers = [nn.Dense(gpu, i, o, w_init=nn.HeNormal(gpu, i, seed=10 + k)) for k, (i, o) in enumerate(dims)]
    net = nn.Sequence([layers[0], nn.ReLU(), layers[1], nn.Sigmoid(), layers[2]], nn.SoftmaxCrossEntropyLoss())
    X, Y = A(gpu, x), A(gpu, y)
    b1t = b2t = 1.0
    for step in range(3):
        params = [(l.w, l.b) for l in layers]
        pre = [[np.asarray(p.value).copy() for p in wb] for wb in params]
        mom = [[(np.asarray(p.opt_state.m).copy(), np.asarray(p.opt_state.v).copy()) for p in wb] for wb in params]
        _, loss = net.train(X, Y)
        Ws = [w for w, _ in pre]
        bs = [b for _, b in pre]
        h1 = O.relu(O.batch_affine(Ws[0], bs[0], x))
        h2 = O.sigmoid(O.batch_affine(Ws[1], bs[1], h1))
        o = O.batch_affine(Ws[2], bs[2], h2)
        sm = O.softmax(o)
        L = O.batch_mean(O.cross_entropy_rows(sm, y))
        assert abs(float(np.asarray(loss)) - float(L)) <= 1e-5 * max(1.0, abs(float(L)))
        d3 = O.loss_grad("sce", sm, y, float(f32(1 / B)))
        dW3, db3, dx3 = O.dense_backward_reference(h2, d3, Ws[2])
        d2 = O.sigmoid_backward(h2, dx3)
        dW2, db2, dx2 = O.dense_backward_reference(h1, d2, Ws[1])
        d1 = O.relu_backward(h1, dx2)
        dW1, db1, _ = O.dense_backward_reference(x, d1, Ws[0])
        b1t *= 0.9
        b2t *= 0.999
        for k, ref in enumerate(((dW1, db1), (dW2, db2), (dW3, db3))):
            for j, (p, g_ref) in enumerate(zip(params[k], ref)):
                g = np.asarray(p.grad)
                scale = max(float(np.abs(g_ref).max()), 1e-12)
                assert np.abs(g - g_ref).max() <= 1e-4 * scale, (step, k, j)
                w_new, m_new, v_new = O.adam(pre[k][j], g, mom[k][j][0], mom[k][j][1],
                                             1e-3, 0.9, 0.999, 1e-8, b1t, b2t)
                bits_equal(p.value, w_new)
                bits_equal(p.opt_state.m, m_new)
                bits_equal(p.opt_state.v, v_new)
