"""Pins of the xfibres_pred restatement (oracle/pred_oracle.py,
xfibres_pred.cc:153-190): closed forms at b = 0 and for perpendicular sticks,
and agreement with the signal model of the oracle's chain (an independent
implementation of the same ball-and-sticks forms, runmcmc_kernels.cu:67-95)."""
import numpy as np
import pytest

from fdt_amd import synth
from oracle import pred_oracle as P


def unit(th, ph):
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 0)


def rand_voxels(n, K, g):
    th = g.uniform(0, np.pi, (K, n))
    ph = g.uniform(-np.pi, np.pi, (K, n))
    f = g.dirichlet(np.ones(K + 1), n).T[:K] * 0.9
    return dict(S0=g.uniform(800, 1200, n), d=g.uniform(0.8e-3, 2e-3, n), dstd=g.uniform(2e-4, 8e-4, n),
                R=g.uniform(0.1, 0.2, n), f0=g.uniform(0.0, 0.05, n), f=f, th=th, ph=ph,
                dyads=np.stack([unit(th[k], ph[k]) for k in range(K)], 0))


def test_b0_gives_S0_and_perpendicular_sticks_one():
    g = np.random.default_rng(1)
    v = rand_voxels(20, 2, g)
    for model in (1, 2, 3):
        for f0 in (None, v["f0"]):
            out = P.predict(np.zeros(1), np.array([[1.0], [0.0], [0.0]]), v["S0"], v["d"], v["f"], v["dyads"], model,
                            v["dstd"], v["R"], f0)
            np.testing.assert_allclose(out[0], v["S0"], rtol=1e-6)
    # model 1, sticks along z, gradient along x: stick terms are f, the ball exp(-b d)
    dz = np.zeros((2, 3, 20))
    dz[:, 2, :] = 1.0
    out = P.predict(np.array([1000.0]), np.array([[1.0], [0.0], [0.0]]), v["S0"], v["d"], v["f"], dz, 1)
    fs = v["f"].sum(0)
    np.testing.assert_allclose(out[0], v["S0"] * (fs + (1 - fs) * np.exp(-1000 * v["d"])), rtol=1e-6)


@pytest.mark.parametrize("model,use_f0", [(1, False), (1, True), (2, False), (2, True), (3, False), (3, True)])
def test_matches_the_chain_signal_model(orc, model, use_f0):
    g = np.random.default_rng(model * 10 + use_f0)
    K, n = 3, 25
    v = rand_voxels(n, K, g)
    bvals, bvecs = synth.acquisition(3, [(1000.0, 20), (3000.0, 20)], g)
    out = P.predict(bvals, bvecs, v["S0"], v["d"], v["f"], v["dyads"], model, v["dstd"], v["R"],
                    v["f0"] if use_f0 else None)
    o = orc.opts(nfibres=K, model=model, f0=int(use_f0))
    for i in range(n):
        st = [v["S0"][i], v["d"][i], v["dstd"][i] if model >= 2 else 0.0, v["R"][i] if model == 3 else 0.0,
              v["f0"][i] if use_f0 else 0.0]
        for k in range(K):
            st += [v["th"][k, i], v["ph"][k, i], v["f"][k, i]]
        want = orc.mcmc_predict(o, bvals, bvecs, np.array(st))
        np.testing.assert_allclose(out[:, i], want, rtol=2e-6)
