"""--gradnonlin on the GPU against the oracle: the per-voxel corrected
acquisitions (correct_bvals_bvecs, CUDA/xfibres_gpu.cu:151-170) bit for bit,
the LM pipeline in FP64 mode, the chains draw for draw and in posterior, and
the xfibres_gpu driver reading a grad_dev part file."""
import os
import subprocess

import numpy as np
import pytest

from fdt_amd import synth

pytestmark = pytest.mark.gpu


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def engine(fdt, v, grad, **opts):
    x = fdt.Xfibres(**opts)
    x.set_acquisition(v["bvals"], v["bvecs"])
    x.load_data(v["data"])
    if grad is not None:
        x.load_grad_dev(grad)
    return x


@pytest.mark.parametrize("name", ["c0", "c1"])
def test_corrected_acquisition_bit_exact(orc, fdt, name):
    n = 40
    v = synth.gradnonlin_voxels(name, n, sd=0.05)
    bv, bc = f32(v["bvals"]), f32(v["bvecs"])
    x = engine(fdt, v, v["grad"], nfibres=2, model=1)
    try:
        for i in range(n):
            gb, gc = x.voxel_acquisition(i)
            wb, wc = orc.correct_acq(bv, bc, v["grad"][i].astype(np.float64))
            np.testing.assert_array_equal(gb.astype(np.float64), wb)
            np.testing.assert_array_equal(gc.astype(np.float64), wc)
        # the tile can change between calls: the table follows it
        x.set_mcmc_tile(32)
        gb, gc = x.voxel_acquisition(n - 1)
        np.testing.assert_array_equal(gb.astype(np.float64), wb)
        x.load_grad_dev(None)
        gb, _ = x.voxel_acquisition(0)
        np.testing.assert_array_equal(gb, v["bvals"].astype(np.float32))
    finally:
        x.close()


@pytest.mark.parametrize("model,nonlin,f0,rician", [(1, 0, 0, 0), (1, 1, 0, 0), (2, 0, 1, 0), (3, 0, 0, 1)])
def test_fit_pipeline_fp64_matches_oracle(orc, fdt, model, nonlin, f0, rician):
    from test_gpu_lm import lm_errors
    name = "c0" if model == 1 else "c1"
    K = 2 if model == 1 else 3
    v = synth.gradnonlin_voxels(name, 48, sd=0.04)
    o = orc.opts(nfibres=K, model=model, nonlin=nonlin, f0=f0, rician=rician)
    want, tau_w, _ = orc.fit(o, f32(v["bvals"]), f32(v["bvecs"]), v["data"].astype(np.float64), grad=v["grad"])
    plain, _, _ = orc.fit(o, f32(v["bvals"]), f32(v["bvecs"]), v["data"].astype(np.float64))
    x = engine(fdt, v, v["grad"], nfibres=K, model=model, nonlin=nonlin, f0=f0, rician=rician, lm_fp64=1)
    try:
        x.fit()
        got, tau_g = x.get_params()
    finally:
        x.close()
    kind = 2 if model >= 2 else (0 if nonlin else 1)
    e = lm_errors(got.astype(np.float64), want, kind, K, bool(f0))
    e_plain = lm_errors(plain, want, kind, K, bool(f0))
    print(f"model={model} nonlin={nonlin} f0={f0} rician={rician}: frac<1e-4 {np.mean(e < 1e-4):.3f} "
          f"(the plain fit: {np.mean(e_plain < 1e-4):.3f})")
    assert np.mean(e < 1e-4) >= 0.8
    assert np.mean(e_plain < 1e-4) < 0.2   # the deviations matter at this size
    if rician:
        agree = e < 1e-5
        np.testing.assert_allclose(tau_g[agree], tau_w[agree], rtol=1e-4)


def run_chains(fdt, v, grad, params, tau, vox_offset=0, **opts):
    x = engine(fdt, v, grad, **opts)
    try:
        x.set_params(params.astype(np.float32), None if tau is None else tau.astype(np.float32))
        x.mcmc(vox_offset)
        x.finish()
        return x.outputs()
    finally:
        x.close()


def oracle_chains(orc, v, grad, params, tau, prec=64, **opts):
    o = orc.opts(**opts)
    s = orc.mcmc(o, f32(v["bvals"]), f32(v["bvecs"]), v["data"].astype(np.float64), f32(params),
                 None if tau is None else f32(tau), grad=grad, prec=prec)
    fin = orc.finish(o, s)
    fin.update(s)
    return fin


@pytest.mark.parametrize("model,K,rician", [(1, 2, 0), (2, 3, 0), (3, 3, 1)])
def test_short_chain_tracks_oracle_draw_for_draw(orc, fdt, model, K, rician):
    name = "c0" if model == 1 else "c1"
    v = synth.gradnonlin_voxels(name, 64, sd=0.04)
    base = dict(nfibres=K, model=model, rician=rician)
    params, tau, _ = orc.fit(orc.opts(**base), f32(v["bvals"]), f32(v["bvecs"]), v["data"].astype(np.float64),
                             grad=v["grad"])
    run = dict(base, nburn=0, njumps=50, sampleevery=25)
    g = run_chains(fdt, v, v["grad"], params, tau if rician else None, **run)
    o = oracle_chains(orc, v, v["grad"], params, tau if rician else None, **run)
    same = np.all(np.abs(g["f"][:, 0, :] - o["f"][:, 0, :]) < 1e-3, axis=0) & \
        (np.abs(g["d"][0] / o["d"][0] - 1) < 1e-3)
    print(f"model {model} rician {rician}: first sample agrees on {same.mean():.3f} of voxels")
    assert same.mean() >= 0.75


def test_posterior_parity_with_oracle(orc, fdt):
    """full chains with per-voxel corrected acquisitions against the FP64
    oracle, held to the oracle's own floors (its FP32 restatement with the
    same draws; itself with another seed), as the shared-acquisition chains"""
    import posterior as PS
    v = synth.gradnonlin_voxels("c1", 256, sd=0.04)
    base = dict(nfibres=3, model=2)
    params, tau, _ = orc.fit(orc.opts(**base), f32(v["bvals"]), f32(v["bvecs"]), v["data"].astype(np.float64),
                             grad=v["grad"])
    run = dict(base, nburn=1000, njumps=1250, sampleevery=25)
    g = run_chains(fdt, v, v["grad"], params, None, **run)
    o = oracle_chains(orc, v, v["grad"], params, None, **run)
    o32 = oracle_chains(orc, v, v["grad"], params, None, prec=32, **run)
    ob = oracle_chains(orc, v, v["grad"], params, None, seed=12345, **run)
    track, prec, floor, cross = PS.stats(g, o), PS.stats(o32, o), PS.stats(ob, o), PS.stats(g, ob)
    print(f"--gradnonlin: GPU vs oracle {PS.fmt(track)} | FP32 vs FP64 {PS.fmt(prec)} | other seed: "
          f"oracle {PS.fmt(floor)}, GPU {PS.fmt(cross)}")
    for m in ("ard", "df", "dyad"):
        n = max(track["n2"], 1) if m == "dyad" else track["n"]
        assert track[m] >= prec[m] - PS.tol(prec[m], n), (m, track[m], prec[m])
        assert cross[m] >= floor[m] - PS.tol(floor[m], n), (m, cross[m], floor[m])


def test_errors_and_reset(fdt):
    v = synth.gradnonlin_voxels("c0", 20)
    x = fdt.Xfibres(nfibres=2, model=1, nburn=10, njumps=25, sampleevery=25)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        x.load_grad_dev(v["grad"][:10])
        with pytest.raises(RuntimeError, match="gradient deviations for 10 voxels"):
            x.fit()
        x.load_grad_dev(np.ascontiguousarray(v["grad"].T), layout=fdt.LAYOUT_VOL_MAJOR)
        x.fit()
        a, _ = x.get_params()
        x.load_grad_dev(v["grad"])
        x.fit()
        b, _ = x.get_params()
        np.testing.assert_array_equal(a, b)          # [9][nvox] and [nvox][9] agree
        x.load_grad_dev(None)
        x.fit()
        c, _ = x.get_params()
        assert not np.array_equal(a, c)
    finally:
        x.close()


def test_driver_reads_grad_dev_part(fdt, tmp_path):
    nvox, nparts, part = 60, 2, 1
    v = synth.gradnonlin_voxels("c0", nvox)
    subj = tmp_path / "subj"
    (tmp_path / "subj.bedpostX" / "logs" / "monitor").mkdir(parents=True)
    (tmp_path / "subj.bedpostX" / "logs" / "logs_gpu").mkdir(parents=True)
    subj.mkdir()
    np.savetxt(tmp_path / "bvals", v["bvals"][None, :])
    np.savetxt(tmp_path / "bvecs", v["bvecs"])
    lo = part * (nvox // nparts)
    np.ascontiguousarray(v["data"][lo:].T).astype(np.float64).tofile(tmp_path / "data_1")
    np.ascontiguousarray(v["grad"][lo:].T).astype(np.float64).tofile(tmp_path / "grad_dev_1")
    logdir = tmp_path / "parts"
    cmd = [fdt.CLI_PATH, f"--data={tmp_path / 'data_1'}", f"--mask={tmp_path / 'mask'}", "-b", str(tmp_path / "bvals"),
           "-r", str(tmp_path / "bvecs"), "--forcedir", f"--logdir={logdir}", "--nf=2", "--fudge=1", "--bi=50",
           "--nj=100", "--se=25", "--model=1", "--cnonlinear", f"--gradnonlin={tmp_path / 'grad_dev_1'}",
           str(subj), str(part), str(nparts), str(nvox)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    th1 = np.fromfile(logdir / "th1samplesJ", dtype=np.float64).reshape(4, nvox - lo)
    ref = fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"][lo:], nfibres=2, model=1, nburn=50, njumps=100,
                          sampleevery=25, vox_offset=lo, grad=v["grad"][lo:])
    np.testing.assert_array_equal(th1, ref["th"][0].astype(np.float64))
    plain = fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"][lo:], nfibres=2, model=1, nburn=50, njumps=100,
                            sampleevery=25, vox_offset=lo)
    assert not np.array_equal(th1, plain["th"][0].astype(np.float64))
