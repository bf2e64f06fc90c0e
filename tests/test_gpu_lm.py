"""GPU parity of the LM initialisation against the FP64 oracle (through the C ABI).

Three arithmetic modes (fdt_opts.lm_fp64): 1 runs the reference LM algorithm
in double precision and must land where the oracle lands; 0 is the
reference's own FP32 arithmetic, held to the error the FP32 restatement
itself shows against FP64; 2 (production) takes every accept/reject and stop
decision on an FP64 cost with FP32 normal equations.  Parameters compared:
S0, d, d_std (relative), volume fractions (absolute), and the orientation of
every fibre with f > 0.05 (angle)."""
import numpy as np
import pytest

from fdt_amd import synth

pytestmark = pytest.mark.gpu

TWO_PI = 0.636619772


def f2x(f):
    return np.tan(f / TWO_PI)


def unit(th, ph):
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)


def lm_errors(a, b, kind, K, f0):
    """per-voxel max error: scalars relative, fractions absolute, angles (rad) for f>0.05"""
    fb = 3 if kind == 2 else 2
    sc = [0, 1] + ([2] if kind == 2 else [])
    e = np.abs(a[:, sc] - b[:, sc]) / np.maximum(np.abs(b[:, sc]), 1e-12)
    fi = [fb + 3 * k for k in range(K)] + ([a.shape[1] - 1] if f0 else [])
    e = np.maximum(e.max(1), np.abs(a[:, fi] - b[:, fi]).max(1))
    for k in range(K):
        xa = unit(a[:, fb + 3 * k + 1], a[:, fb + 3 * k + 2])
        xb = unit(b[:, fb + 3 * k + 1], b[:, fb + 3 * k + 2])
        ang = np.arccos(np.clip(np.abs((xa * xb).sum(1)), 0, 1))
        e = np.maximum(e, np.where(b[:, fb + 3 * k] > 0.05, ang, 0.0))
    return e


def stage_inputs(orc, kind, K, f0, G, nvox=48):
    name = "c0" if kind < 2 else "c1"
    v = synth.voxels(name, nvox)
    y = v["data"].astype(np.float64)
    if kind < 2:
        st = orc.start(kind, K, f0, v["bvals"], v["bvecs"], y)
    else:
        s0 = orc.start(orc.K_SINGLE, K, f0, v["bvals"], v["bvecs"], y)
        s1, _, _ = orc.lm(orc.K_SINGLE, K, f0, 0, 0.0, v["bvals"], v["bvecs"], y, s0)
        np1 = 2 + 3 * K + int(f0)
        st = np.zeros((nvox, np1 + 1))
        st[:, 0], st[:, 1], st[:, 2] = s1[:, 0], 1.0, s1[:, 1]
        for k in range(K):
            st[:, 3 + 3 * k] = f2x(s1[:, 2 + 3 * k])
            st[:, 4 + 3 * k] = s1[:, 3 + 3 * k]
            st[:, 5 + 3 * k] = s1[:, 4 + 3 * k]
        if f0:
            st[:, -1] = f2x(s1[:, np1 - 1])
    return v, y, st


CASES = [(0, 0), (1, 0), (2, 0), (2, 2)]


@pytest.mark.parametrize("kind,G", CASES)
@pytest.mark.parametrize("K", [1, 2, 3])
@pytest.mark.parametrize("f0", [False, True])
def test_lm_stage_fp64_matches_oracle(orc, fdt, kind, G, K, f0):
    v, y, st = stage_inputs(orc, kind, K, f0, G)
    R = 0.13
    want, it_o, _ = orc.lm(kind, K, f0, G, R, v["bvals"], v["bvecs"], y, st)
    x = fdt.Xfibres(nfibres=K, model=1 if kind < 2 else 2, lm_fp64=1)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        got, it_g = x.lm_stage(kind, K, f0, G, R, st.astype(np.float32))
    finally:
        x.close()
    # start values enter as float32: compare against the oracle started from the same float32 starts
    want32, it_o32, _ = orc.lm(kind, K, f0, G, R, v["bvals"], v["bvecs"], y, st.astype(np.float32).astype(np.float64))
    e = lm_errors(got.astype(np.float64), want32, kind, K, f0)
    print(f"kind={kind} G={G} K={K} f0={f0}: frac<1e-5 {np.mean(e < 1e-5):.3f} "
          f"iters equal {np.mean(it_g == it_o32):.3f} median err {np.median(e):.2e}")
    assert np.mean(e < 1e-5) >= 0.85
    assert np.median(e) < 1e-6


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("kind,G", CASES)
def test_lm_stage_fp32_error_is_fp32_rounding(orc, fdt, kind, G, mode):
    """mode 0: FP32 model, Jacobian and cost -- no worse than the reference's
    own FP32 arithmetic (the oracle's FP32 restatement).  mode 2 (production):
    FP32 Jacobian and normal equations with FP64 parameters and cost -- within
    1e-4 of the FP64 oracle on >= 90% of voxels."""
    K, f0 = 2, False
    v, y, st = stage_inputs(orc, kind, K, f0, G, nvox=256)
    R = 0.13
    st32 = st.astype(np.float32).astype(np.float64)
    want, _, _ = orc.lm(kind, K, f0, G, R, v["bvals"], v["bvecs"], y, st32)
    ref32, _, _ = orc.lm(kind, K, f0, G, R, v["bvals"], v["bvecs"], y, st32, prec=32)
    x = fdt.Xfibres(nfibres=K, model=1 if kind < 2 else 2, lm_fp64=mode)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        got, _ = x.lm_stage(kind, K, f0, G, R, st.astype(np.float32))
    finally:
        x.close()
    e_gpu = lm_errors(got.astype(np.float64), want, kind, K, f0)
    e_ref = lm_errors(ref32, want, kind, K, f0)
    print(f"kind={kind} G={G} mode={mode}: engine within 1e-4 {np.mean(e_gpu < 1e-4):.3f} median "
          f"{np.median(e_gpu):.2e} p90 {np.percentile(e_gpu, 90):.2e}; FP32 restatement within 1e-4 "
          f"{np.mean(e_ref < 1e-4):.3f} median {np.median(e_ref):.2e} p90 {np.percentile(e_ref, 90):.2e}")
    if mode == 0:
        # no worse than the reference's own FP32 arithmetic (x3 slack on the median)
        assert np.median(e_gpu) <= 3 * np.median(e_ref) + 1e-6
        assert np.median(e_gpu) < 1e-3
    else:
        assert np.mean(e_gpu < 1e-4) >= 0.9


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("K", [1, 3])
def test_start_values_match_oracle(orc, fdt, kind, K):
    v = synth.voxels("c0", 64)
    y = v["data"].astype(np.float64)
    want = orc.start(kind, K, False, v["bvals"], v["bvecs"], y)
    x = fdt.Xfibres(nfibres=K, model=1, lm_fp64=1)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        got = x.lm_start(kind, K, False).astype(np.float64)
    finally:
        x.close()
    # the tensor's minor eigenvectors are ill-defined when eigenvalues nearly tie;
    # compare S0, d and the principal direction everywhere, all values on >= 90%
    np.testing.assert_allclose(got[:, 0], want[:, 0], rtol=1e-5)
    np.testing.assert_allclose(got[:, 1], want[:, 1], rtol=1e-4)
    x1g, x1w = unit(got[:, 3], got[:, 4]), unit(want[:, 3], want[:, 4])
    assert np.all(np.abs((x1g * x1w).sum(1)) > 1 - 1e-8)
    close = np.all(np.abs(got - want) <= 1e-4 * np.maximum(np.abs(want), 1.0), axis=1)
    assert close.mean() >= 0.9


def final_cost(orc, o, v, params, model, K, f0):
    """sum of squared residuals (FP64) of finished fits [S0, d, (d_std), (f, th, ph) x K, (f0)]: the LM
    objective of the last stage (model 3: zeppelin sticks at R = Rmean)"""
    add = 1 if model >= 2 else 0
    out = np.zeros(params.shape[0])
    for i, p in enumerate(params.astype(np.float64)):
        st = np.zeros(5 + 3 * K)
        st[0], st[1] = abs(p[0]), p[1]
        st[2] = p[2] if add else 0.0
        st[3] = o.Rmean
        st[4] = p[-1] if f0 else 0.0
        for k in range(K):
            st[5 + 3 * k], st[6 + 3 * k], st[7 + 3 * k] = p[3 + add + 3 * k], p[4 + add + 3 * k], p[2 + add + 3 * k]
        pred = orc.mcmc_predict(o, v["bvals"], v["bvecs"], st)
        out[i] = ((v["data"][i].astype(np.float64) - pred) ** 2).sum()
    return out


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("model,nonlin,f0,rician", [
    (1, 0, 0, 0), (1, 1, 0, 0), (1, 0, 1, 0), (2, 0, 0, 0), (2, 0, 1, 0), (3, 0, 0, 1), (1, 0, 0, 1)])
def test_fit_pipeline_matches_oracle(orc, fdt, model, nonlin, f0, rician, mode):
    """The whole fit (tensor start, stage 1, stage 2, --f0 refit, tau) against
    the FP64 oracle.  The fit is a chaotic function of its arithmetic: the
    oracle against itself with the data perturbed by 1e-8 relative (below the
    float32 resolution of the data) already misses 1e-4 on a few percent of
    voxels (different stopping points in flat valleys, different local minima,
    different --f0 refit decisions).  FP64 mode is held to that floor;
    production mode (FP32 normal equations, FP64 parameters and cost) to it
    within a margin.  For the misses the final costs are compared: the GPU
    must not land on systematically worse optima."""
    name = "c0" if model == 1 else "c1"
    K = 2 if model == 1 else 3
    nvox = 384
    v = synth.voxels(name, nvox)
    y = v["data"].astype(np.float64)
    o = orc.opts(nfibres=K, model=model, nonlin=nonlin, f0=f0, rician=rician)
    want, tau_w, _ = orc.fit(o, v["bvals"], v["bvecs"], y)
    yp = y * (1 + 1e-8 * np.random.default_rng(1).standard_normal(y.shape))
    wantp, _, _ = orc.fit(o, v["bvals"], v["bvecs"], yp)
    x = fdt.Xfibres(nfibres=K, model=model, nonlin=nonlin, f0=f0, rician=rician, lm_fp64=mode)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        x.fit()
        got, tau_g = x.get_params()
    finally:
        x.close()
    kind = 2 if model >= 2 else (0 if nonlin else 1)
    e = lm_errors(got.astype(np.float64), want, kind, K, bool(f0))
    ef = lm_errors(wantp, want, kind, K, bool(f0))
    within, floor = np.mean(e < 1e-4), np.mean(ef < 1e-4)
    miss = e >= 1e-4
    c_g = final_cost(orc, o, v, got[miss], model, K, f0)
    c_o = final_cost(orc, o, v, want[miss], model, K, f0)
    rc = (c_g - c_o) / c_o
    lower = np.mean(rc <= 0) if miss.any() else 1.0
    print(f"model={model} nonlin={nonlin} f0={f0} rician={rician} mode={mode}: within 1e-4 {within:.3f} "
          f"(oracle vs perturbed oracle {floor:.3f}); {miss.sum()} misses: final cost within 1e-6 of the "
          f"oracle's {np.mean(np.abs(rc) <= 1e-6) if miss.any() else 1.0:.3f}, GPU cost lower or equal {lower:.3f}, "
          f"median |rel cost diff| {np.median(np.abs(rc)) if miss.any() else 0.0:.2e}")
    tol = 3 * np.sqrt(max(floor * (1 - floor), 0.01) / nvox)
    # production mode: the FP32 Hessian sets different steps along the flat
    # directions; held to the floor within 0.15 (--f0 fits run in FP64)
    margin = 0.0 if (mode == 1 or f0) else 0.15
    assert within >= floor - tol - margin
    if miss.sum() >= 20:
        assert lower >= 0.3   # misses are not systematically worse optima
    if rician:
        agree = e < 1e-5
        np.testing.assert_allclose(tau_g[agree], tau_w[agree], rtol=1e-4 if mode == 1 else 1e-3)


def test_rician_nonpositive_samples_replaced(orc, fdt):
    v = synth.voxels("c0", 16)
    d = v["data"].copy()
    d[:, 20] = 0.0
    d[:, 21] = -5.0
    o = orc.opts(nfibres=2, model=1, rician=1)
    want, tau_w, _ = orc.fit(o, v["bvals"], v["bvecs"], d.astype(np.float64))
    x = fdt.Xfibres(nfibres=2, model=1, rician=1, lm_fp64=1)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(d)
        x.fit()
        got, tau_g = x.get_params()
    finally:
        x.close()
    assert np.all(np.isfinite(got)) and np.all(np.isfinite(tau_g))
    e = lm_errors(got.astype(np.float64), want, 1, 2, False)
    assert np.mean(e < 1e-4) >= 0.8
