"""Behavioural pins of the oracle's LM fit and MCMC (restating
CUDA/levenberg_marquardt.cu, CUDA/xfibres_gpu.cu fit(), CUDA/runmcmc_kernels.cu
and CUDA/samples.cu): noise-free recovery, stationarity, the --f0 refit rule,
Rician tau, chain invariants, determinism and the finishing step."""
import numpy as np
import pytest

from fdt_amd import synth


def unit(th, ph):
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)


def match_angle_deg(p, v, K, fb):
    """angle between each true fibre with f>0.05 and the closest fitted fibre"""
    out = []
    for i in range(p.shape[0]):
        fx = unit(p[i, [fb + 3 * k + 1 for k in range(K)]], p[i, [fb + 3 * k + 2 for k in range(K)]])
        for k in range(K):
            if v["f"][i, k] > 0.05:
                c = np.abs(fx @ v["dirs"][i, k]).max()
                out.append(np.degrees(np.arccos(min(1.0, c))))
    return np.array(out)


def test_lm_noise_free_model1_recovers_truth(orc):
    # LM is a local optimiser started from the tensor fit: on noise-free data
    # nearly every voxel (>= 95%) lands on the truth.
    v = synth.voxels("c0", 64)
    clean = v["clean"].astype(np.float64)
    for nonlin in (0, 1):
        o = orc.opts(nfibres=2, model=1, nonlin=nonlin)
        p, _, _ = orc.fit(o, v["bvals"], v["bvecs"], clean)
        fs = np.sort(p[:, [2, 5]], axis=1)[:, ::-1]
        ft = np.sort(v["f"][:, :2], axis=1)[:, ::-1]
        ok = (np.abs(p[:, 0] / v["S0"] - 1) < 1e-4) & (np.abs(p[:, 1] / v["d"] - 1) < 1e-3) \
            & (np.abs(fs[:, 0] - ft[:, 0]) < 2e-3) & ((ft[:, 1] < 0.05) | (np.abs(fs[:, 1] - ft[:, 1]) < 2e-3))
        assert ok.mean() >= 0.95, (nonlin, ok.mean())
        assert np.percentile(match_angle_deg(p, v, 2, 2), 90) < 0.5


def test_lm_noise_free_model2_recovers_truth(orc):
    v = synth.voxels("c1", 48)
    o = orc.opts(nfibres=3, model=2)
    p, _, _ = orc.fit(o, v["bvals"], v["bvecs"], v["clean"].astype(np.float64))
    # PVM_multi starts from shape a = 1 and uses Levenberg damping H + lambda*I
    # (diffmodels.cc:1375-1393, CUDA/PVM_multi.cu:520-528): on a third of the
    # voxels it stops well short of the optimum, as the reference does, so
    # only S0, the orientations and the median diffusivity are pinned here.
    assert (np.abs(p[:, 0] / v["S0"] - 1) < 1e-2).mean() >= 0.9
    assert np.median(np.abs(p[:, 1] / v["d"] - 1)) < 5e-2
    assert np.median(match_angle_deg(p, v, 3, 3)) < 1.0


def test_lm_stationary_on_noisy_data(orc):
    # At the LM end point the gradient 2 J^T r vanishes relative to its terms,
    # and the cost is no worse than at the start values.
    v = synth.voxels("c0", 32)
    y = v["data"].astype(np.float64)
    st = orc.start(orc.K_SINGLE_C, 2, False, v["bvals"], v["bvecs"], y)
    out, it, cost = orc.lm(orc.K_SINGLE_C, 2, False, 0, 0.0, v["bvals"], v["bvecs"], y, st)
    for i in range(y.shape[0]):
        c0 = orc.model_eval(orc.K_SINGLE_C, 2, False, 0, 0.0, v["bvals"], v["bvecs"], st[i], y[i])[3]
        assert cost[i] <= c0 * (1 + 1e-12)
    # gradient at the end: recompute from the transformed end point
    i = 0
    lam = np.sqrt(out[i, 1])
    fs = out[i, [2, 5]]
    b1 = np.arcsin(np.sqrt(fs[0]))
    b2 = np.arcsin(np.sqrt(min(1.0, fs[1] / (1 - fs[0]))))
    pt = np.array([out[i, 0], lam, b1, out[i, 3], out[i, 4], b2, out[i, 6], out[i, 7]])
    pred, Jg, _, _ = orc.model_eval(orc.K_SINGLE_C, 2, False, 0, 0.0, v["bvals"], v["bvecs"], pt)
    r = pred - y[i]
    g = 2 * Jg.T @ r
    scale = 2 * np.abs(Jg).T @ np.abs(r)
    assert np.all(np.abs(g) <= 1e-3 * scale + 1e-9)


def test_f0_implausible_refit(orc):
    # a voxel dominated by an unattenuated compartment: the --f0 fit is implausible
    # (f0 > 0.4 or f1 < 0.05) and is redone without f0, f0 set to 0.001
    v = synth.voxels("c0", 8)
    y = v["clean"].astype(np.float64)
    y = 0.35 * y + 0.65 * v["S0"][:, None]
    o = orc.opts(nfibres=2, model=1, f0=1)
    p, _, _ = orc.fit(o, v["bvals"], v["bvecs"], y)
    assert np.all(p[:, -1] == pytest.approx(0.001))


def test_rician_tau_is_inverse_residual_variance(orc):
    v = synth.voxels("c0", 16)
    y = v["data"].astype(np.float64)
    o = orc.opts(nfibres=2, model=1, rician=1)
    p, tau, _ = orc.fit(o, v["bvals"], v["bvecs"], y)
    b, g = v["bvals"], v["bvecs"]
    for i in range(4):
        S0, d = p[i, 0], p[i, 1]
        pred = (1 - p[i, 2] - p[i, 5]) * np.exp(-b * d)
        for k in range(2):
            x = unit(p[i, 3 + 3 * k], p[i, 4 + 3 * k])
            pred += p[i, 2 + 3 * k] * np.exp(-b * d * (x @ g) ** 2)
        r = y[i] - S0 * pred
        assert tau[i] == pytest.approx(1.0 / np.var(r, ddof=1), rel=1e-6)


@pytest.fixture(scope="module")
def chain(orc):
    v = synth.voxels("c0", 24)
    y = v["data"].astype(np.float64)
    o = orc.opts(nfibres=2, model=1, nburn=300, njumps=500, sampleevery=25)
    p, tau, _ = orc.fit(o, v["bvals"], v["bvecs"], y)
    s = orc.mcmc(o, v["bvals"], v["bvecs"], y, p, tau)
    return v, o, y, p, tau, s


def test_mcmc_invariants(chain):
    v, o, y, p, tau, s = chain
    assert s["th"].shape == (2, 20, 24)
    f = s["f"]
    assert np.all(f > 0) and np.all(f < 1)
    assert np.all(f.sum(0) <= 1.0 + 1e-6)
    assert np.all(s["d"] >= 0) and np.all(s["d"] <= 0.005)
    assert np.all(s["S0"] >= 0)
    assert np.isfinite(s["energy"]).all()


def test_mcmc_seed_determinism(orc, chain):
    v, o, y, p, tau, s = chain
    s2 = orc.mcmc(o, v["bvals"], v["bvecs"], y, p, tau)
    for k in ("th", "ph", "f", "S0", "d"):
        np.testing.assert_array_equal(s[k], s2[k])
    # the draws are keyed by the global voxel index: a shifted offset changes the chain
    s3 = orc.mcmc(o, v["bvals"], v["bvecs"], y, p, tau, vox_offset=1)
    assert not np.array_equal(s["f"], s3["f"])
    # ... and splitting the voxels with matching offsets does not
    a = orc.mcmc(o, v["bvals"], v["bvecs"], y[:10], p[:10], tau[:10], vox_offset=0)
    b = orc.mcmc(o, v["bvals"], v["bvecs"], y[10:], p[10:], tau[10:], vox_offset=10)
    np.testing.assert_array_equal(np.concatenate([a["f"], b["f"]], axis=2), s["f"])


def test_mcmc_high_snr_posterior_near_truth(orc):
    v = synth.voxels("c0", 16)
    # SNR ~ 200: posterior means close to the truth
    rng = np.random.default_rng(3)
    y = v["clean"] + rng.standard_normal(v["clean"].shape) * (v["S0"][:, None] / 200)
    o = orc.opts(nfibres=2, model=1, nburn=500, njumps=500, sampleevery=10)
    p, tau, _ = orc.fit(o, v["bvals"], v["bvecs"], y)
    s = orc.mcmc(o, v["bvals"], v["bvecs"], y, p, tau)
    fin = orc.finish(o, s)
    # the chain starts from the LM fit, which (rarely) sits in a local minimum:
    # pin >= 90% of the voxels
    assert (np.abs(fin["mean_S0"] / v["S0"] - 1) < 2e-2).mean() >= 0.9
    assert (np.abs(fin["mean_d"] / v["d"] - 1) < 1e-1).mean() >= 0.9
    f1 = fin["mean_f"][0]
    assert (np.abs(f1 - np.sort(v["f"][:, :2], 1)[:, 1]) < 0.05).mean() >= 0.9
    ang = []
    for i in range(16):
        c = abs(fin["dyads"][0, :, i] @ v["dirs"][i, np.argmax(v["f"][i])])
        ang.append(np.degrees(np.arccos(min(1.0, c))))
    assert np.median(ang) < 3.0


def test_finish_sort_means_dyads(orc, chain):
    v, o, y, p, tau, s = chain
    raw = {k: (None if val is None else val.copy()) for k, val in s.items()}
    fin = orc.finish(o, s)   # permutes th/ph/f in place
    ns, nv = 20, 24
    for i in range(nv):
        sums = raw["f"][:, :, i].sum(1)
        order = sorted(range(2), key=lambda k: (sums[k], k), reverse=True)
        for j, src in enumerate(order):
            np.testing.assert_array_equal(s["f"][j, :, i], raw["f"][src, :, i])
            np.testing.assert_array_equal(s["th"][j, :, i], raw["th"][src, :, i])
            x = unit(raw["th"][src, :, i].astype(np.float64), raw["ph"][src, :, i].astype(np.float64))
            T = x.T @ x / ns
            w, V = np.linalg.eigh(T)
            assert abs(abs(V[:, -1] @ fin["dyads"][j, :, i]) - 1) < 1e-9
            assert fin["disp"][j, i] == pytest.approx(1 - w[-1], abs=1e-12)
            assert fin["mean_f"][j, i] == pytest.approx(raw["f"][src, :, i].astype(np.float64).mean(), rel=1e-12)
    np.testing.assert_allclose(fin["mean_d"], raw["d"].astype(np.float64).mean(0), rtol=1e-12)
