"""Golden fixtures of the FP64 oracle (tests/golden/oracle_golden.npz, written
by tests/golden/make_golden.py): the oracle must still reproduce them (CPU),
and the HIP engine must land on them (GPU) -- LM stages in FP64 mode within
1e-5, the fit pipeline within 1e-4, the MCMC chain draw for draw."""
import os

import numpy as np
import pytest

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_golden.npz"))
LM_CASES = sorted({k[:-len("_start")] for k in G.files if k.startswith("lm_") and k.endswith("_start")})
FIT_CASES = sorted({k[:-len("_params")] for k in G.files if k.startswith("fit_") and k.endswith("_params")})


def acq(name):
    return G[f"acq_{name}_bvals"], G[f"acq_{name}_bvecs"]


def lm_case(key):
    _, name, k, g, n, f = key.split("_")
    return name, int(k[1:]), int(g[1:]), int(n[1:]), bool(int(f[1:]))


def fit_case(key):
    parts = key.split("_")
    return parts[1], int(parts[2][1:]), int(parts[3][1:]), parts[-1] == "ric"


@pytest.mark.parametrize("key", LM_CASES)
def test_oracle_reproduces_lm_fixture(orc, key):
    name, kind, g, K, f0 = lm_case(key)
    bv, bc = acq(name)
    out, it, cost = orc.lm(kind, K, f0, g, 0.13, bv, bc, G[key + "_data"].astype(np.float64), G[key + "_start"])
    np.testing.assert_allclose(out, G[key + "_out"], rtol=1e-10, atol=1e-14)
    np.testing.assert_array_equal(it, G[key + "_iters"])
    np.testing.assert_allclose(cost, G[key + "_cost"], rtol=1e-10)


@pytest.mark.parametrize("key", FIT_CASES)
def test_oracle_reproduces_fit_fixture(orc, key):
    name, model, K, ric = fit_case(key)
    bv, bc = acq(name)
    o = orc.opts(nfibres=K, model=model, rician=int(ric))
    p, tau, _ = orc.fit(o, bv, bc, G[key + "_data"].astype(np.float64))
    np.testing.assert_allclose(p, G[key + "_params"], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(tau, G[key + "_tau"], rtol=1e-10, atol=1e-14)


def test_oracle_reproduces_mcmc_fixture(orc):
    bv, bc = acq("c0")
    o = orc.opts(nfibres=2, model=1, nburn=50, njumps=100, sampleevery=25)
    s = orc.mcmc(o, bv, bc, G["mcmc_c0_data"].astype(np.float64), G["mcmc_c0_params"], G["mcmc_c0_tau"])
    for k in ("th", "ph", "f", "S0", "d", "energy"):
        np.testing.assert_allclose(np.asarray(s[k], dtype=np.float64), G[f"mcmc_c0_{k}"], rtol=1e-12, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("key", LM_CASES)
def test_engine_lm_stage_on_fixture(fdt, key):
    name, kind, g, K, f0 = lm_case(key)
    bv, bc = acq(name)
    x = fdt.Xfibres(nfibres=K, model=1 if kind < 2 else 2, lm_fp64=1)
    try:
        x.set_acquisition(bv, bc)
        x.load_data(G[key + "_data"])
        got, _ = x.lm_stage(kind, K, f0, g, 0.13, G[key + "_start"].astype(np.float32))
    finally:
        x.close()
    want = G[key + "_out"]
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
    ok = np.all(rel <= 1e-5, axis=1)
    print(f"{key}: {ok.sum()}/{ok.size} voxels within 1e-5")
    assert ok.mean() >= 0.75


@pytest.mark.gpu
@pytest.mark.parametrize("key", FIT_CASES)
def test_engine_fit_on_fixture(fdt, key):
    name, model, K, ric = fit_case(key)
    bv, bc = acq(name)
    x = fdt.Xfibres(nfibres=K, model=model, rician=int(ric), lm_fp64=1)
    try:
        x.set_acquisition(bv, bc)
        x.load_data(G[key + "_data"])
        x.fit()
        got, tau = x.get_params()
    finally:
        x.close()
    want = G[key + "_params"]
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
    ok = np.all(rel <= 1e-4, axis=1)
    print(f"{key}: {ok.sum()}/{ok.size} voxels within 1e-4")
    assert ok.mean() >= 0.75
    if ric:
        np.testing.assert_allclose(tau[ok], G[key + "_tau"][ok], rtol=1e-3)


@pytest.mark.gpu
def test_engine_mcmc_on_fixture(fdt):
    bv, bc = acq("c0")
    x = fdt.Xfibres(nfibres=2, model=1, nburn=50, njumps=100, sampleevery=25)
    try:
        x.set_acquisition(bv, bc)
        x.load_data(G["mcmc_c0_data"])
        x.set_params(G["mcmc_c0_params"].astype(np.float32), G["mcmc_c0_tau"].astype(np.float32))
        x.mcmc(0)
        x.finish()
        out = x.outputs()
    finally:
        x.close()
    # raw (unsorted) oracle samples vs the engine's sorted ones: compare the sorted order of both
    f_o = G["mcmc_c0_f"]
    order = np.argsort(-f_o.sum(1), axis=0, kind="stable")  # [K][nvox] fibre ranks
    f_os = np.take_along_axis(f_o, order[:, None, :], axis=0)
    same = np.all(np.abs(out["f"][:, 0, :] - f_os[:, 0, :]) < 1e-3, axis=0)
    print(f"first sample identical on {same.sum()}/{same.size} voxels")
    assert same.mean() >= 0.75
