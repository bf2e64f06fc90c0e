"""Chain edge cases against the oracle: acquisitions whose length leaves
several padding slots in a lane (N = 50, 100, 200), tiles larger than the
acquisition needs (bit-identical to the smallest tile), chains whose angles
are far outside one revolution (as width adaptation drives the angles of
unsupported fibres), where sin/cos must be reduced over the full range as
the reference does (CUDA/runmcmc_kernels.cu:340, :1165, :1177-1182), and an
acquisition with more distinct b values than a b-shell table holds, where
the chains and the LM cost evaluate the isotropic signal per volume."""
import numpy as np
import pytest

from fdt_amd import synth

pytestmark = pytest.mark.gpu


def subset(v, n):
    """The first n volumes of a workload's acquisition and data."""
    return dict(bvals=np.ascontiguousarray(v["bvals"][:n]), bvecs=np.ascontiguousarray(v["bvecs"][:, :n]),
                data=np.ascontiguousarray(v["data"][:, :n]))


def run_engine(fdt, v, params, tau, vpl=0, **opts):
    x = fdt.Xfibres(**opts)
    try:
        if vpl:
            x.set_mcmc_tile(vpl)
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        x.set_params(params.astype(np.float32), None if tau is None else tau.astype(np.float32))
        x.mcmc(0)
        x.finish()
        return x.outputs()
    finally:
        x.close()


def first_sample_agreement(g, o):
    same = np.all(np.abs(g["f"][:, 0, :] - o["f"][:, 0, :]) < 1e-3, axis=0) & \
        (np.abs(g["d"][0] / o["d"][0] - 1) < 1e-3)
    return same.mean()


@pytest.mark.parametrize("n,model,K", [(50, 1, 2), (100, 2, 3), (200, 2, 3)])
def test_padded_lanes_track_oracle_draw_for_draw(orc, fdt, n, model, K):
    """N=50 fills 3 full slots and 2 lanes of a fourth in the 5-slot tile;
    N=100 and N=200 leave 3 and 5 slots of padding in their tiles."""
    v = subset(synth.voxels("c1", 64), n)
    base = dict(nfibres=K, model=model)
    params, tau, _ = orc.fit(orc.opts(**base), v["bvals"], v["bvecs"], v["data"].astype(np.float64))
    run = dict(base, nburn=0, njumps=50, sampleevery=25)
    g = run_engine(fdt, v, params, tau, **run)
    o = orc.mcmc(orc.opts(**run), v["bvals"], v["bvecs"], v["data"].astype(np.float64),
                 params.astype(np.float32).astype(np.float64), tau.astype(np.float32).astype(np.float64))
    orc.finish(orc.opts(**run), o)
    agree = first_sample_agreement(g, o)
    print(f"N={n} model {model} K={K}: first sample agrees on {agree:.3f} of voxels")
    assert agree >= 0.75


@pytest.mark.parametrize("name,model,K,rician,tiles", [
    ("c2_33", 1, 1, 0, (3, 5, 10)),
    ("c2_69", 2, 2, 0, (5, 10)),
    ("c2_33", 2, 2, 1, (3, 10)),
])
def test_larger_tiles_are_bit_identical(orc, fdt, name, model, K, rician, tiles):
    v = synth.voxels(name, 40)
    base = dict(nfibres=K, model=model, rician=rician)
    params, tau, _ = orc.fit(orc.opts(**base), v["bvals"], v["bvecs"], v["data"].astype(np.float64))
    run = dict(base, nburn=60, njumps=100, sampleevery=25)
    outs = [run_engine(fdt, v, params, tau if rician else None, vpl=t, **run) for t in tiles]
    for t, o in zip(tiles[1:], outs[1:]):
        for k in ("th", "ph", "f", "S0", "d", "mean_S0", "dyads"):
            np.testing.assert_array_equal(o[k], outs[0][k], err_msg=f"tile {t} vs {tiles[0]}: {k}")


@pytest.mark.parametrize("model,K", [(1, 2), (2, 3)])
def test_wide_angle_chains_track_fp32_oracle(orc, fdt, model, K):
    """Chains started with every fibre angle near 1.2e9 rad.  The initial
    proposal widths (0.2) are far below the float spacing there (128), so both
    chains keep those angles for the first sample and every likelihood is
    evaluated at directions that exist only through the full-range reduction
    of sin/cos.  Compared with the oracle's FP32 restatement, whose angles are
    floats as well."""
    name = "c0" if model == 1 else "c1"
    v = synth.voxels(name, 64)
    base = dict(nfibres=K, model=model)
    params, tau, _ = orc.fit(orc.opts(**base), v["bvals"], v["bvecs"], v["data"].astype(np.float64))
    add = 1 if model >= 2 else 0
    p = params.astype(np.float32).copy()
    for k in range(K):
        for j, big in ((3 + 3 * k + add, 1.234567e9), (4 + 3 * k + add, -2.345678e9)):
            p[:, j] = np.float32(p[:, j] + big)
    run = dict(base, nburn=0, njumps=50, sampleevery=25)
    g = run_engine(fdt, v, p, tau, **run)
    o = orc.mcmc(orc.opts(**run), v["bvals"], v["bvecs"], v["data"].astype(np.float64), p.astype(np.float64),
                 tau.astype(np.float32).astype(np.float64), prec=32)
    orc.finish(orc.opts(**run), o)
    # the angles did not move (proposals round back to the start value)
    np.testing.assert_array_equal(np.sort(g["th"][:, 0, :], axis=0), np.sort(p[:, [3 + 3 * k + add for k in range(K)]].T, axis=0))
    agree = first_sample_agreement(g, o)
    print(f"model {model}: first sample at angles ~1e9 agrees with the FP32 oracle on {agree:.3f} of voxels")
    assert agree >= 0.75


def jittered_b_voxels(nvox=64, seed=5):
    """c1 voxels acquired with every non-zero b value jittered by U(-30, 30)
    s/mm^2 (271 distinct values; a b-shell table holds 16), signal and Rician
    noise (SNR 25) generated for that acquisition."""
    v = synth.voxels("c1", nvox)
    g = np.random.default_rng(seed)
    b = v["bvals"].astype(np.float64).copy()
    nz = b > 0
    b[nz] += g.uniform(-30.0, 30.0, int(nz.sum()))
    b = b.astype(np.float32).astype(np.float64)
    sig = synth.clean_signal(2, b, v["bvecs"].astype(np.float64), v["S0"], v["d"], v["dstd"], v["f"], v["dirs"])
    sigma = (v["S0"] / 25.0)[:, None]
    n1, n2 = g.standard_normal(sig.shape) * sigma, g.standard_normal(sig.shape) * sigma
    data = np.sqrt((sig + n1) ** 2 + n2 ** 2).astype(np.float32)
    return dict(bvals=b.astype(np.float32), bvecs=np.ascontiguousarray(v["bvecs"], dtype=np.float32), data=data)


def test_many_b_values_lm_and_chains_track_oracle(orc, fdt):
    from test_gpu_lm import lm_errors
    v = jittered_b_voxels()
    assert np.unique(v["bvals"]).size > 16
    b64, g64, y64 = (v["bvals"].astype(np.float64), v["bvecs"].astype(np.float64), v["data"].astype(np.float64))
    base = dict(nfibres=3, model=2)
    want, _, _ = orc.fit(orc.opts(**base), b64, g64, y64)
    for mode, bar in ((1, 0.8), (2, 0.75)):
        x = fdt.Xfibres(lm_fp64=mode, **base)
        try:
            x.set_acquisition(v["bvals"], v["bvecs"])
            x.load_data(v["data"])
            x.fit()
            got, _ = x.get_params()
        finally:
            x.close()
        e = lm_errors(got.astype(np.float64), want, 2, 3, False)
        print(f"{np.unique(v['bvals']).size} distinct b values, LM mode {mode}: within 1e-4 {np.mean(e < 1e-4):.3f}")
        assert np.mean(e < 1e-4) >= bar
    run = dict(base, nburn=0, njumps=50, sampleevery=25)
    g = run_engine(fdt, v, want, None, **run)
    o = orc.mcmc(orc.opts(**run), b64, g64, y64, want.astype(np.float32).astype(np.float64), None)
    orc.finish(orc.opts(**run), o)
    agree = first_sample_agreement(g, o)
    print(f"chains without a b-shell table: first sample agrees on {agree:.3f} of voxels")
    assert agree >= 0.75
