"""GPU MCMC, finishing step and driver, against the FP64 oracle.

Both sides draw from Philox4x32-10 with the same (seed, voxel, jump, slot)
counters, so from identical LM parameters the chains agree draw for draw
until a near-tie accept/reject decision flips; the statistical bar of the
hot path (posterior-mean dyads within 2 degrees on 95% of voxels with 2+
supported fibres, |mean f| within 0.02, same supported-fibre counts on 98%)
is measured on full-length chains beside the oracle's own floors."""
import os
import subprocess

import numpy as np
import pytest

from fdt_amd import synth
import posterior as PS

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_engine(fdt, v, data, params, tau, vox_offset=0, **opts):
    x = fdt.Xfibres(**opts)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(data)
        x.set_params(params.astype(np.float32), None if tau is None else tau.astype(np.float32))
        x.mcmc(vox_offset)
        x.finish()
        return x.outputs()
    finally:
        x.close()


def oracle_chain(orc, v, data, params, tau, prec=64, **opts):
    o = orc.opts(**opts)
    s = orc.mcmc(o, v["bvals"], v["bvecs"], data.astype(np.float64), params.astype(np.float32).astype(np.float64),
                 None if tau is None else tau.astype(np.float32).astype(np.float64), prec=prec)
    fin = orc.finish(o, s)
    fin.update(s)
    return fin


def unit(th, ph):
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)


@pytest.mark.parametrize("model,K", [(1, 2), (2, 3)])
def test_short_chain_tracks_oracle_draw_for_draw(orc, fdt, model, K):
    name = "c0" if model == 1 else "c1"
    v = synth.voxels(name, 64)
    base = dict(nfibres=K, model=model)
    params, tau, _ = orc.fit(orc.opts(**base), v["bvals"], v["bvecs"], v["data"].astype(np.float64))
    run = dict(base, nburn=0, njumps=50, sampleevery=25)
    g = run_engine(fdt, v, v["data"], params, tau, **run)
    o = oracle_chain(orc, v, v["data"], params, tau, **run)
    # same first sample (after 25 sweeps) on most voxels
    same = np.all(np.abs(g["f"][:, 0, :] - o["f"][:, 0, :]) < 1e-3, axis=0) & \
        (np.abs(g["d"][0] / o["d"][0] - 1) < 1e-3)
    print(f"model {model}: first sample agrees on {same.mean():.3f} of voxels")
    assert same.mean() >= 0.75


@pytest.mark.parametrize("name,model,K,nvox", [("c0", 1, 2, 1024), ("c1", 2, 3, 768)])
def test_posterior_parity_with_oracle(orc, fdt, name, model, K, nvox):
    """Full-length chains (1,000 burn-in + 1,250 recorded sweeps) on the GPU
    against the FP64 oracle.  The posterior parity bar (ARD counts equal on 98% of
    voxels, |mean f| within 0.02 and dyads within 2 degrees on 95%) is not met
    by the oracle against itself: with another seed (Monte Carlo noise of 50
    recorded samples), or in the reference's FP32 arithmetic with the same
    draws.  The GPU chain is therefore held to those floors: with the same
    seed it tracks the FP64 oracle at least as well as the FP32 restatement
    does, and against the oracle under another seed it agrees as well as the
    oracle agrees with itself; where a floor meets the bar, the bar itself is
    asserted."""
    v = synth.voxels(name, nvox)
    base = dict(nfibres=K, model=model)
    params, tau, _ = orc.fit(orc.opts(**base), v["bvals"], v["bvecs"], v["data"].astype(np.float64))
    run = dict(base, nburn=1000, njumps=1250, sampleevery=25)
    seed_b = 12345
    g = run_engine(fdt, v, v["data"], params, tau, **run)
    o = oracle_chain(orc, v, v["data"], params, tau, **run)
    o32 = oracle_chain(orc, v, v["data"], params, tau, prec=32, **run)
    ob = oracle_chain(orc, v, v["data"], params, tau, seed=seed_b, **run)
    track, prec, floor, cross = PS.stats(g, o), PS.stats(o32, o), PS.stats(ob, o), PS.stats(g, ob)
    for lab, s in (("GPU vs oracle, same seed", track), ("oracle FP32 vs FP64, same seed", prec),
                   ("oracle vs oracle, other seed", floor), ("GPU vs oracle, other seed", cross)):
        print(f"{name} model {model} K={K}, {nvox} voxels, {lab}: {PS.fmt(s)}")
    for m in ("ard", "df", "dyad"):
        n = track["n2"] if m == "dyad" else track["n"]
        assert track[m] >= prec[m] - PS.tol(prec[m], n), (m, track[m], prec[m])
        assert cross[m] >= floor[m] - PS.tol(floor[m], n), (m, cross[m], floor[m])
        if prec[m] >= PS.BAR[m]:
            assert track[m] >= PS.BAR[m] - PS.tol(PS.BAR[m], n), (m, track[m])
    for k in ("mean_S0", "mean_d", "mean_dstd"):
        if k in g and np.any(o[k]):
            print(f"{k}: median relative difference to the oracle: GPU {np.median(np.abs(g[k] / o[k] - 1)):.2e} "
                  f"(same seed), {np.median(np.abs(g[k] / ob[k] - 1)):.2e} (other seed); oracle other seed "
                  f"{np.median(np.abs(ob[k] / o[k] - 1)):.2e}")
    s0 = np.abs(g["mean_S0"] / o["mean_S0"] - 1)
    s0f = np.abs(ob["mean_S0"] / o["mean_S0"] - 1)
    print(f"mean S0 within 2% of the oracle: GPU {np.mean(s0 <= 0.02):.3f}, oracle other seed {np.mean(s0f <= 0.02):.3f}")
    assert np.mean(s0 <= 0.02) >= np.mean(s0f <= 0.02) - PS.tol(np.mean(s0f <= 0.02), len(s0))


def test_seed_determinism_and_partition_independence(orc, fdt):
    v = synth.voxels("c0", 40)
    base = dict(nfibres=2, model=1, nburn=50, njumps=100, sampleevery=25)
    params, tau, _ = orc.fit(orc.opts(nfibres=2, model=1), v["bvals"], v["bvecs"], v["data"].astype(np.float64))
    a = run_engine(fdt, v, v["data"], params, tau, **base)
    b = run_engine(fdt, v, v["data"], params, tau, **base)
    for k in ("th", "ph", "f", "mean_S0", "mean_d", "dyads"):
        np.testing.assert_array_equal(a[k], b[k])
    lo = run_engine(fdt, v, v["data"][:17], params[:17], tau[:17], vox_offset=0, **base)
    hi = run_engine(fdt, v, v["data"][17:], params[17:], tau[17:], vox_offset=17, **base)
    np.testing.assert_array_equal(np.concatenate([lo["f"], hi["f"]], axis=2), a["f"])
    np.testing.assert_array_equal(np.concatenate([lo["mean_d"], hi["mean_d"]]), a["mean_d"])
    c = run_engine(fdt, v, v["data"], params, tau, seed=1234, **base)
    assert not np.array_equal(c["f"], a["f"])


@pytest.mark.parametrize("name,model,K,opts", [
    ("c2_33", 1, 1, {}),
    ("c0", 1, 2, {"f0": 1, "ardf0": 1}),
    ("c0", 1, 2, {"noard": 1}),
    ("c0", 1, 2, {"allard": 1}),
    ("c2_145", 2, 3, {}),
    ("c1", 3, 3, {"rician": 1}),
    ("c1", 3, 2, {"f0": 1, "Rfudge": 1.0}),
])
def test_option_matrix_runs_within_bounds(fdt, name, model, K, opts):
    kw = {"noard": "no_ard", "allard": "all_ard", "Rfudge": "R_prior_fudge"}
    o = {kw.get(k, k): val for k, val in opts.items()}
    for nvox in (1, 17):
        v = synth.voxels(name, nvox)
        out = fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"], nfibres=K, model=model, nburn=100, njumps=200,
                              sampleevery=25, **o)
        f = out["f"]
        assert f.shape == (K, 8, nvox)
        assert np.all(np.isfinite(f)) and np.all(f > 0) and np.all(f < 1)
        assert np.all(f.sum(0) <= 1 + 1e-5)
        assert np.all((out["d"] >= 0) & (out["d"] <= 0.005))
        assert np.all(out["S0"] >= 0)
        np.testing.assert_allclose(np.linalg.norm(out["dyads"], axis=1), 1.0, rtol=1e-5)
        if model >= 2:
            up = 0.01 if model == 2 else 0.004
            assert np.all((out["dstd"] > 0) & (out["dstd"] <= up))
        if model == 3:
            assert np.all(out["R"] > 0) and np.all(out["R"] <= 1)
        if opts.get("f0"):
            assert np.all((out["f0"] > 0) & (out["f0"] < 1))
        if opts.get("rician"):
            assert np.all(out["tau"] > 0)


def test_finish_outputs_consistent_with_samples(orc, fdt):
    v = synth.voxels("c0", 33)
    out = fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"], nfibres=2, model=1, nburn=100, njumps=250,
                          sampleevery=25)
    f = out["f"].astype(np.float64)
    sums = f.sum(1)
    assert np.all(sums[0] >= sums[1])      # sorted by descending sum of f
    np.testing.assert_allclose(out["mean_f"], f.mean(1), rtol=1e-5)
    np.testing.assert_allclose(out["mean_d"], out["d"].astype(np.float64).mean(0), rtol=1e-5)
    for i in range(33):
        for k in range(2):
            x = unit(out["th"][k, :, i].astype(np.float64), out["ph"][k, :, i].astype(np.float64))
            w, V = np.linalg.eigh(x.T @ x / x.shape[0])
            assert abs(abs(V[:, -1] @ out["dyads"][k, :, i]) - 1) < 1e-5
            assert out["disp"][k, i] == pytest.approx(1 - w[-1], abs=1e-5)


def test_empty_volume(fdt):
    v = synth.voxels("c0", 1)
    x = fdt.Xfibres(nfibres=2, model=1)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(np.zeros((0, 69), np.float32))
        x.fit()
        x.mcmc(0)
        x.finish()
        out = x.outputs()
        assert out["f"].shape == (2, 50, 0)
    finally:
        x.close()


def test_volume_major_layout_matches_voxel_major(fdt):
    v = synth.voxels("c0", 50)
    kw = dict(nfibres=2, model=1, nburn=20, njumps=50, sampleevery=25)
    a = fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"], **kw)
    b = fdt.run_xfibres(v["bvals"], v["bvecs"], np.ascontiguousarray(v["data"].T), layout=fdt.LAYOUT_VOL_MAJOR, **kw)
    for k in ("f", "th", "mean_S0", "params"):
        np.testing.assert_array_equal(a[k], b[k])


def test_driver_end_to_end(fdt, tmp_path):
    """xfibres_gpu on a bedpostx-style part: <name>J files in --logdir hold the
    library's outputs for the part's global voxel range."""
    nvox, nparts, part = 100, 2, 1
    v = synth.voxels("c0", nvox)
    subj = tmp_path / "subj"
    (tmp_path / "subj.bedpostX" / "logs" / "monitor").mkdir(parents=True)
    (tmp_path / "subj.bedpostX" / "logs" / "logs_gpu").mkdir(parents=True)
    subj.mkdir()
    np.savetxt(tmp_path / "bvals", v["bvals"][None, :])
    np.savetxt(tmp_path / "bvecs", v["bvecs"])
    size = nvox // nparts
    lo = part * size
    part_data = np.ascontiguousarray(v["data"][lo:].T).astype(np.float64)   # [N][nvox_part] Real
    part_data.tofile(tmp_path / "data_1")
    logdir = tmp_path / "parts"
    cmd = [fdt.CLI_PATH, f"--data={tmp_path / 'data_1'}", f"--mask={tmp_path / 'mask'}", "-b", str(tmp_path / "bvals"),
           "-r", str(tmp_path / "bvecs"), "--forcedir", f"--logdir={logdir}", "--nf=2", "--fudge=1", "--bi=50",
           "--nj=100", "--se=25", "--model=1", "--cnonlinear", str(subj), str(part), str(nparts), str(nvox)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Part processed in" in r.stdout
    ns = 4
    th1 = np.fromfile(logdir / "th1samplesJ", dtype=np.float64).reshape(ns, nvox - lo)
    mS0 = np.fromfile(logdir / "mean_S0samplesJ", dtype=np.float64)
    assert mS0.shape == (nvox - lo,)
    for name in ("ph1samplesJ", "f1samplesJ", "th2samplesJ", "ph2samplesJ", "f2samplesJ", "mean_dsamplesJ"):
        assert (logdir / name).exists()
    assert (tmp_path / "subj.bedpostX" / "logs" / "monitor" / "1").read_bytes() == b"done"
    ref = fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"][lo:], nfibres=2, model=1, nburn=50, njumps=100,
                          sampleevery=25, vox_offset=lo)
    np.testing.assert_array_equal(th1, ref["th"][0].astype(np.float64))
    np.testing.assert_array_equal(mS0, ref["mean_S0"].astype(np.float64))
