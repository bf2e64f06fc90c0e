"""Pins of the oracle's LM forward models and Jacobians (CUDA/PVM_single.cu,
CUDA/PVM_single_c.cu, CUDA/PVM_multi.cu restated in oracle/xfibres_oracle.cpp):
closed forms, the gamma-to-exponential limit, and central finite differences
of the forward model against the Jacobian columns (with the reference's two
documented deviations from the exact derivative)."""
import numpy as np
import pytest

from fdt_amd import synth

TWO_PI = 0.636619772


def f2x(f):
    return np.tan(f / TWO_PI)


def acq(n=40, seed=3):
    g = np.random.default_rng(seed)
    v = g.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    b = np.concatenate([[0.0, 0.0], np.full(n - 2, 1000.0)])
    b[n // 2:] = 2500.0
    return b, np.ascontiguousarray(v.T)


def params(kind, K, f0, rng):
    p = []
    p.append(rng.uniform(800, 1200))
    if kind == 0:
        p.append(rng.uniform(1e-3, 2e-3))
    elif kind == 1:
        p.append(np.sqrt(rng.uniform(1e-3, 2e-3)))
    else:
        p += [rng.uniform(2.0, 8.0), rng.uniform(2e-4, 6e-4)]
    for _ in range(K):
        if kind == 1:
            p.append(rng.uniform(0.3, 1.0))   # beta
        else:
            p.append(f2x(rng.uniform(0.1, 0.25)))
        p += [rng.uniform(0.3, 2.8), rng.uniform(-3.0, 3.0)]
    if f0:
        p.append(rng.uniform(0.05, 0.2) if kind == 1 else f2x(0.05))
    return np.array(p)


@pytest.mark.parametrize("kind,G", [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2)])
@pytest.mark.parametrize("K", [1, 2, 3])
@pytest.mark.parametrize("f0", [False, True])
def test_jacobian_finite_differences(orc, kind, G, K, f0):
    b, g = acq()
    rng = np.random.default_rng(100 * kind + 10 * K + G + int(f0))
    p = params(kind, K, f0, rng)
    Rv = 0.13
    pred, Jg, Jh, _ = orc.model_eval(kind, K, f0, G, Rv, b, g, p)
    fb = 3 if kind == 2 else 2
    for j in range(p.size):
        h = 1e-6 * max(abs(p[j]), 1e-4)
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (orc.model_eval(kind, K, f0, G, Rv, b, g, pp)[0] - orc.model_eval(kind, K, f0, G, Rv, b, g, pm)[0]) / (2 * h)
        is_angle = j >= fb and j < fb + 3 * K and (j - fb) % 3 != 0
        is_frac = j >= fb and j < fb + 3 * K and (j - fb) % 3 == 0
        if kind == 1 and f0 and is_frac:
            # PVM_single_c's beta_k columns omit the dependence of f0 = sin^2(beta0)(1 - sum f)
            # on beta_k (CUDA/PVM_single_c.cu:296-306): the exact derivative adds that term.
            continue
        if kind == 2 and G == 2 and is_angle:
            # the G=2 angle derivative carries no factor (g.x) (CUDA/PVM_multi.cu:77-78):
            # J = d pred / d angle divided by (g.x)
            k = (j - fb) // 3
            th, ph = p[fb + 3 * k + 1], p[fb + 3 * k + 2]
            x = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
            dp = x @ g
            ok = np.abs(dp) > 1e-3
            np.testing.assert_allclose(Jg[ok, j] * dp[ok], fd[ok], rtol=2e-5, atol=1e-6 * np.abs(fd).max())
            continue
        np.testing.assert_allclose(Jg[:, j], fd, rtol=2e-5, atol=1e-6 * max(np.abs(fd).max(), 1e-12))
        np.testing.assert_allclose(Jh[:, j], Jg[:, j], rtol=1e-12)


def test_closed_forms_b0_and_alignment(orc):
    # b = 0 gives S0 for every model; a stick along g gives exp(-b d) whatever f;
    # a stick orthogonal to g gives S0((1-f) exp(-b d) + f).
    S0, d, f = 1000.0, 1.5e-3, 0.4
    th, ph = 0.7, 1.1
    x = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    o = np.cross(x, [0.0, 0.0, 1.0])
    o /= np.linalg.norm(o)
    g = np.ascontiguousarray(np.stack([x, x, o]).T)
    b = np.array([0.0, 1000.0, 1000.0])
    for kind, p in [(0, [S0, d, f2x(f), th, ph]), (1, [S0, np.sqrt(d), np.arcsin(np.sqrt(f)), th, ph])]:
        pred = orc.model_eval(kind, 1, False, 0, 0.0, b, g, np.array(p))[0]
        np.testing.assert_allclose(pred, [S0, S0 * np.exp(-1000 * d), S0 * ((1 - f) * np.exp(-1000 * d) + f)], rtol=1e-12)


def test_multi_gamma_limit_is_exponential(orc):
    # PVM_multi with shape a -> inf at fixed mean d = a*b is the exponential model
    b, g = acq()
    d = 1.2e-3
    a = 1e7
    p_multi = np.array([1000.0, a, d / a, f2x(0.3), 0.9, 0.4])
    p_single = np.array([1000.0, d, f2x(0.3), 0.9, 0.4])
    pm = orc.model_eval(2, 1, False, 0, 0.0, b, g, p_multi)[0]
    ps = orc.model_eval(0, 1, False, 0, 0.0, b, g, p_single)[0]
    np.testing.assert_allclose(pm, ps, rtol=1e-5)


def test_zeppelin_r1_is_isotropic(orc):
    # G=2 with R=1: exp(-3 b a b' /(2R+1) ((1-R) dp^2 + R)) = exp(-b d), no direction dependence
    b, g = acq()
    a, bb = 3.0, 5e-4
    p = np.array([1000.0, a, bb, f2x(0.5), 0.3, 2.0])
    pred = orc.model_eval(2, 1, False, 2, 1.0, b, g, p)[0]
    iso = np.exp(-a * np.log(1 + b * bb))
    np.testing.assert_allclose(pred, 1000.0 * (0.5 * iso + 0.5 * np.exp(-b * a * bb)), rtol=1e-9)


def test_mcmc_angle_term_identity():
    # the reference angle term (fibre.h:353-358) equals (g.x)^2 -- the form the engine uses
    rng = np.random.default_rng(7)
    for _ in range(100):
        g = rng.standard_normal(3)
        g /= np.linalg.norm(g)
        alpha = np.arctan2(np.hypot(g[0], g[1]), g[2])
        beta = np.arctan2(g[1], g[0])
        th, ph = rng.uniform(-4, 4, 2)
        cm, cp = np.cos(alpha - th), np.cos(alpha + th)
        ref = (np.cos(ph - beta) * (cm - cp) / 2 + (cm + cp) / 2) ** 2
        x = [np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]
        assert ref == pytest.approx(np.dot(g, x) ** 2, abs=1e-12)


@pytest.mark.parametrize("model", [1, 2, 3])
def test_mcmc_forward_model_closed_forms(orc, model):
    # MCMC signal (CUDA/runmcmc_kernels.cu:67-95): b=0 -> S0; f=0 -> iso only;
    # model 2 with tiny d_std -> model 1; model 3 with R=1 -> direction independent
    S0, d = 1000.0, 1.4e-3
    b, g = acq(30)
    o = orc.opts(nfibres=1, model=model)
    st = np.array([S0, d, 3e-4 if model >= 2 else 0.0, 1.0 if model == 3 else 0.13, 0.0, 0.6, 0.2, 0.0])
    pred = orc.mcmc_predict(o, b, g, st)
    if model == 1:
        iso = np.exp(-d * b)
    else:
        s2 = 3e-4 ** 2
        iso = np.exp(np.log(d / (d + b * s2)) * d * d / s2)
    np.testing.assert_allclose(pred, S0 * iso, rtol=1e-12)
    st[7] = 0.5
    pred = orc.mcmc_predict(o, b, g, st)
    assert pred[0] == pytest.approx(S0)
    if model == 2:
        st[2] = 1e-7   # below 1e-5: exponential forms (fibre.h:351, :914)
        p2 = orc.mcmc_predict(o, b, g, st)
        o1 = orc.opts(nfibres=1, model=1)
        p1 = orc.mcmc_predict(o1, b, g, st)
        np.testing.assert_allclose(p2, p1, rtol=1e-12)
    if model == 3:
        np.testing.assert_allclose(pred, S0 * (0.5 * iso + 0.5 * np.exp(-b * d)), rtol=1e-9)


def test_logI0_against_scipy(orc):
    sp = pytest.importorskip("scipy.special")
    for x in [1e-4, 0.1, 1.0, 3.7, 3.76, 10.0, 100.0, 1e4]:
        want = np.log(sp.i0e(x)) + x
        assert orc.logI0(x) == pytest.approx(want, rel=3e-7, abs=3e-7)
