"""Pins of the CPU restatement of the learned start's network
(oracle/mlp_oracle.py): exact-FP64 gradients against central finite
differences, the sign invariance of the direction loss, the BF16 rounding
against its definition, and GELU against its closed form."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def rand_targets(g, n):
    t = np.zeros((n, M.SG_OUT))
    t[:, 0:3] = g.uniform(0, 0.5, (n, 3))
    for k in range(3):
        x = g.standard_normal((n, 3))
        t[:, 3 + 3 * k:6 + 3 * k] = x / np.linalg.norm(x, axis=1, keepdims=True)
    t[:, 12:14] = g.uniform(0.5, 2.0, (n, 2))
    t[:, 14:] = g.normal(-3, 1, (n, 14))
    t[:, 14 + 3] = 1e30  # an unused slot
    return t


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.3e-5, 65504.0], dtype=np.float32)
    got = M.bf16(x)
    # bf16 keeps 8 significant bits: 1 + 2^-8 is a tie -> even (1.0); 1 + 1.5*2^-8 -> 1 + 2^-7
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0078125
    assert abs(got[3] / -3.3e-5 - 1) < 2 ** -8
    assert got[4] == 65536.0


def test_gelu_closed_form():
    z = np.linspace(-5, 5, 101)
    from scipy.stats import norm
    np.testing.assert_allclose(M.gelu(z), z * norm.cdf(z), rtol=1e-12, atol=1e-15)
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), rtol=1e-7, atol=1e-9)


def test_direction_loss_is_sign_free():
    g = np.random.default_rng(1)
    o = g.standard_normal((5, M.SG_OUT))
    t = rand_targets(g, 5)
    l1, _ = M.loss_grad(o, t)
    o2 = o.copy()
    o2[:, 3:6] *= -1
    l2, _ = M.loss_grad(o2, t)
    assert l1 == pytest.approx(l2, rel=1e-12)
    # p = x (unit) gives zero direction loss
    o3 = o.copy()
    o3[:, 3:12] = t[:, 3:12]
    o3[:, 0:3] = t[:, 0:3]
    o3[:, 12:14] = t[:, 12:14]
    o3[:, 14:] = np.where(t[:, 14:] < 1e29, t[:, 14:], 0.0)
    l3, g3 = M.loss_grad(o3, t)
    assert l3 == pytest.approx(0.0, abs=1e-12)
    np.testing.assert_allclose(g3, 0.0, atol=1e-12)


def test_output_gradient_finite_differences():
    g = np.random.default_rng(2)
    o = g.standard_normal((3, M.SG_OUT))
    t = rand_targets(g, 3)
    _, G = M.loss_grad(o, t)
    h = 1e-6
    for i in range(3):
        for j in range(M.SG_OUT):
            op, om = o.copy(), o.copy()
            op[i, j] += h
            om[i, j] -= h
            fd = (M.loss_grad(op, t)[0] - M.loss_grad(om, t)[0]) / (2 * h)
            assert G[i, j] == pytest.approx(fd, rel=1e-5, abs=1e-8)


def test_weight_gradients_finite_differences():
    g = np.random.default_rng(3)
    n_in, b = 24, 6
    layers = M.init_layers(n_in, seed=4)
    layers = [(W, g.standard_normal(bb.shape) * 0.1) for W, bb in layers]
    x = g.uniform(0.1, 1.0, (b, n_in))
    t = rand_targets(g, b)
    grads, loss = M.backward(x, layers, t, emulate_bf16=False)
    h = 1e-6
    for l in range(4):
        W, bias = layers[l]
        for _ in range(4):
            i, j = g.integers(W.shape[0]), g.integers(W.shape[1])
            Wp, Wm = W.copy(), W.copy()
            Wp[i, j] += h
            Wm[i, j] -= h
            lp = M.loss_grad(M.forward(x, layers[:l] + [(Wp, bias)] + layers[l + 1:], False)[0], t)[0]
            lm = M.loss_grad(M.forward(x, layers[:l] + [(Wm, bias)] + layers[l + 1:], False)[0], t)[0]
            assert grads[l][0][i, j] == pytest.approx((lp - lm) / (2 * h), rel=2e-4, abs=1e-9)
        j = g.integers(bias.shape[0])
        bp, bm = bias.copy(), bias.copy()
        bp[j] += h
        bm[j] -= h
        lp = M.loss_grad(M.forward(x, layers[:l] + [(W, bp)] + layers[l + 1:], False)[0], t)[0]
        lm = M.loss_grad(M.forward(x, layers[:l] + [(W, bm)] + layers[l + 1:], False)[0], t)[0]
        assert grads[l][1][j] == pytest.approx((lp - lm) / (2 * h), rel=2e-4, abs=1e-9)


def test_bf16_emulation_close_to_exact():
    g = np.random.default_rng(5)
    layers = M.init_layers(40, seed=6)
    x = g.uniform(0.1, 1.0, (16, 40))
    t = rand_targets(g, 16)
    ge, _ = M.backward(x, layers, t, emulate_bf16=False)
    gq, _ = M.backward(x, layers, t, emulate_bf16=True)
    for (a, _), (b, _) in zip(ge, gq):
        assert np.linalg.norm(a - b) / np.linalg.norm(a) < 0.05
