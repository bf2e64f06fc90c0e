"""The batched queue (fdt_enqueue_batch / fdt_run_queue): volumes with
different acquisitions, models and fibre counts run in one persistent chain
launch and must give what each volume gives on its own (same tiles, same
seeds and voxel offsets); graph replays must repeat the first run; the
autotuner must leave a working configuration."""
import numpy as np
import pytest

from fdt_amd import synth
import posterior as PS

pytestmark = pytest.mark.gpu

MIX = [("c2_33", 1, 1), ("c2_69", 2, 2), ("c2_145", 2, 3), ("c2_288", 3, 3)]
RUN = dict(nburn=40, njumps=100, sampleevery=25)


def build_queue(fdt, nvox, extra=None):
    q = fdt.Queue()
    vols = []
    for j, (name, model, K) in enumerate(MIX):
        v = synth.voxels(name, nvox[j] if isinstance(nvox, list) else nvox)
        kw = dict(RUN, nfibres=K, model=model)
        if extra and j in extra:
            kw.update(extra[j])
        q.enqueue(v["bvals"], v["bvecs"], v["data"], vox_offset=1000 * j, **kw)
        vols.append((v, kw))
    return q, vols


def test_queue_matches_per_volume_runs(fdt):
    q, vols = build_queue(fdt, [37, 64, 50, 29])
    try:
        q.run()
        for j, (v, kw) in enumerate(vols):
            got = q.outputs(j)
            want = fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"], vox_offset=1000 * j, **kw)
            for k in ("th", "ph", "f", "mean_S0", "mean_d", "dyads"):
                np.testing.assert_array_equal(got[k], want[k], err_msg=f"job {j} {k}")
            p, _ = q.params(j)
            np.testing.assert_array_equal(p, want["params"])
    finally:
        q.close()


def test_graph_replay_repeats_first_run(fdt):
    q, _ = build_queue(fdt, 40)
    try:
        q.run()  # eager + capture
        first = [q.outputs(j) for j in range(len(MIX))]
        q.run()  # graph replay
        q.run()
        for j in range(len(MIX)):
            again = q.outputs(j)
            for k in ("th", "f", "mean_d", "dyads"):
                np.testing.assert_array_equal(again[k], first[j][k])
        assert q.timing()["ms"] > 0
    finally:
        q.close()


def test_autotune_then_run(fdt):
    """after the autotuner has chosen workgroup sizes and tiles, every job equals
    its per-volume run at the smallest tile (padding slots contribute nothing)"""
    q, vols = build_queue(fdt, 48)
    try:
        ms = q.autotune(jumps=20, sample_vox=48)
        t = q.timing()
        print(f"autotune: best {ms:.3f} ms, rows/wg {t['rows_per_wg']}, tiles {t['vpl']}")
        assert ms > 0 and t["rows_per_wg"] in (4, 8, 16)
        for j, (name, model, K) in enumerate(MIX):
            assert 16 * t["vpl"][j] >= vols[j][0]["bvals"].size
        q.run()
        for j, (v, kw) in enumerate(vols):
            got = q.outputs(j)
            want = fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"], vox_offset=1000 * j, **kw)
            for k in ("th", "ph", "f", "mean_S0", "mean_d", "dyads"):
                np.testing.assert_array_equal(got[k], want[k], err_msg=f"job {j} {k}")
    finally:
        q.close()


def test_rows_per_workgroup_choices_agree(fdt):
    outs = []
    for rows in (4, 8, 16):
        q, _ = build_queue(fdt, 33)
        try:
            q.set_tiles(rows_per_wg=rows)
            q.run()
            outs.append([q.outputs(j) for j in range(len(MIX))])
        finally:
            q.close()
    for j in range(len(MIX)):
        for other in outs[1:]:
            np.testing.assert_array_equal(other[j]["th"], outs[0][j]["th"])


def test_rician_job_runs_outside_the_queue_kernel(fdt):
    q, vols = build_queue(fdt, 24, extra={3: dict(rician=1)})
    try:
        q.run()
        v, kw = vols[3]
        got = q.outputs(3)
        want = fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"], vox_offset=3000, **kw)
        np.testing.assert_array_equal(got["th"], want["th"])
        assert np.all(got["tau"] > 0)
    finally:
        q.close()


def test_queue_errors(fdt):
    q = fdt.Queue()
    try:
        with pytest.raises(fdt.FdtError, match="empty"):
            q.run()
        v = synth.voxels("c0", 4)
        with pytest.raises(fdt.FdtError):
            q.enqueue(v["bvals"][:5], v["bvecs"][:, :5], v["data"][:, :5], nfibres=2, model=1)
        with pytest.raises(fdt.FdtError):
            q.set_tiles(rows_per_wg=5)
    finally:
        q.close()


def test_gradnonlin_job_in_a_batch(fdt):
    # a --gradnonlin volume among plain ones: its chains run in their own launch,
    # every job still equals its per-volume run, and the graph replay repeats it
    q, vols = build_queue(fdt, 36)
    try:
        v = synth.gradnonlin_voxels("c0", 36, sd=0.04)
        j = q.enqueue(v["bvals"], v["bvecs"], v["data"], vox_offset=5000, nfibres=2, model=1, nburn=20, njumps=50,
                      sampleevery=25)
        q.set_grad_dev(j, v["grad"])
        q.run()
        first = q.outputs(j)
        want = fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"], vox_offset=5000, grad=v["grad"], nfibres=2, model=1,
                               nburn=20, njumps=50, sampleevery=25)
        for k in ("th", "ph", "f", "mean_S0", "dyads"):
            np.testing.assert_array_equal(first[k], want[k], err_msg=k)
        for jj, (vv, kw) in enumerate(vols):
            ref = fdt.run_xfibres(vv["bvals"], vv["bvecs"], vv["data"], vox_offset=1000 * jj, **kw)
            np.testing.assert_array_equal(q.outputs(jj)["th"], ref["th"])
        q.run()  # replay
        np.testing.assert_array_equal(q.outputs(j)["th"], first["th"])
    finally:
        q.close()


def test_tensor_start_job_in_a_batch(fdt):
    # a CPU-style tensor start (cpu_init + nospat) inside a batch equals its per-volume run;
    # neighbour initialisation (sequential wavefronts) is refused in a batch
    q, vols = build_queue(fdt, 30)
    try:
        v = synth.voxels("c0", 40)
        kw = dict(nfibres=2, model=1, nburn=20, njumps=50, sampleevery=25, cpu_init=1, nospat=1)
        j = q.enqueue(v["bvals"], v["bvecs"], v["data"], vox_offset=7000, **kw)
        with pytest.raises(fdt.FdtError, match="no place in a batch"):
            q.enqueue(v["bvals"], v["bvecs"], v["data"], **dict(kw, nospat=0))
        q.run()
        got = q.outputs(j)
        x = fdt.Xfibres(**kw)
        try:
            x.set_acquisition(v["bvals"], v["bvecs"])
            x.load_data(v["data"])
            x.run(7000)
            x.sync()
            want = x.outputs()
        finally:
            x.close()
        for k in ("th", "ph", "f", "mean_S0", "dyads"):
            np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    finally:
        q.close()


def oracle_chain(orc, v, params, tau, vox_offset, prec=64, **run):
    o = orc.opts(**run)
    smp = orc.mcmc(o, v["bvals"], v["bvecs"], v["data"].astype(np.float64), params.astype(np.float64),
                   None if tau is None else tau.astype(np.float64), vox_offset=vox_offset, prec=prec)
    fin = orc.finish(o, smp)
    fin.update(smp)
    return fin


def test_queue_job_types_track_oracle_draw_for_draw(orc, fdt):
    """every job type of the mixed workload (N 33/69/145/288, models 1/2/2/3,
    nf 1/2/3/3) in one batch: from the queue's own LM start, the first recorded
    sample matches the FP64 oracle's chain draw for draw on most voxels"""
    run = dict(nburn=0, njumps=50, sampleevery=25)
    q = fdt.Queue()
    vols = []
    try:
        for j, (name, model, K) in enumerate(MIX):
            v = synth.voxels(name, 64, sub=7)
            q.enqueue(v["bvals"], v["bvecs"], v["data"], vox_offset=1000 * j, nfibres=K, model=model, **run)
            vols.append((v, model, K))
        q.run()
        for j, (v, model, K) in enumerate(vols):
            g = q.outputs(j)
            params, _ = q.params(j)
            o = oracle_chain(orc, v, params, None, 1000 * j, nfibres=K, model=model, **run)
            same = np.all(np.abs(g["f"][:, 0, :] - o["f"][:, 0, :]) < 1e-3, axis=0) & \
                (np.abs(g["d"][0] / o["d"][0] - 1) < 1e-3)
            print(f"job {j} ({MIX[j]}): first sample agrees on {same.mean():.3f} of voxels")
            assert same.mean() >= 0.75
    finally:
        q.close()


def test_queue_job_types_posterior_vs_oracle(orc, fdt):
    """full chains (1,000 + 1,250 sweeps) of every job type in one batch
    against the FP64 oracle from the same start, held to the oracle's own
    floors: the FP32 restatement with the same draws, and the oracle with
    another seed"""
    run = dict(nburn=1000, njumps=1250, sampleevery=25)
    q = fdt.Queue()
    vols = []
    try:
        for j, (name, model, K) in enumerate(MIX):
            v = synth.voxels(name, 96, sub=8)
            q.enqueue(v["bvals"], v["bvecs"], v["data"], vox_offset=1000 * j, nfibres=K, model=model, **run)
            vols.append((v, model, K))
        q.run()
        for j, (v, model, K) in enumerate(vols):
            g = q.outputs(j)
            params, _ = q.params(j)
            kw = dict(nfibres=K, model=model, **run)
            o = oracle_chain(orc, v, params, None, 1000 * j, **kw)
            o32 = oracle_chain(orc, v, params, None, 1000 * j, prec=32, **kw)
            ob = oracle_chain(orc, v, params, None, 1000 * j, seed=12345, **kw)
            track, prec, floor, cross = PS.stats(g, o), PS.stats(o32, o), PS.stats(ob, o), PS.stats(g, ob)
            print(f"job {j} {MIX[j]}: GPU vs oracle {PS.fmt(track)} | FP32 vs FP64 {PS.fmt(prec)} | "
                  f"other seed: oracle {PS.fmt(floor)}, GPU {PS.fmt(cross)}")
            for m in ("ard", "df", "dyad"):
                n = max(track["n2"], 1) if m == "dyad" else track["n"]
                assert track[m] >= prec[m] - PS.tol(prec[m], n), (j, m, track[m], prec[m])
                assert cross[m] >= floor[m] - PS.tol(floor[m], n), (j, m, cross[m], floor[m])
            s0 = np.abs(g["mean_S0"] / o["mean_S0"] - 1)
            s0f = np.abs(ob["mean_S0"] / o["mean_S0"] - 1)
            assert np.mean(s0 <= 0.02) >= np.mean(s0f <= 0.02) - PS.tol(np.mean(s0f <= 0.02), len(s0)), (j, s0.max())
    finally:
        q.close()


def test_full_size_mixed_queue(fdt):
    """the mixed workload at full size: four volumes of 60,000 voxels in one
    batch with the default chains; every output within its bounds"""
    q = fdt.Queue()
    try:
        for j, (name, model, K) in enumerate(MIX):
            v = synth.voxels(name, 60000, with_truth=False)
            q.enqueue(v["bvals"], v["bvecs"], v["data"], vox_offset=100000 * j, nfibres=K, model=model)
        del v
        q.run()
        for j, (name, model, K) in enumerate(MIX):
            out = q.outputs(j)
            f = out["f"]
            assert f.shape == (K, 50, 60000)
            assert np.all(np.isfinite(f)) and np.all((f > 0) & (f < 1)), j
            assert np.all(f.sum(0) <= 1 + 1e-5), j
            np.testing.assert_allclose(np.linalg.norm(out["dyads"], axis=1), 1.0, rtol=1e-5)
            assert np.all((out["d"] >= 0) & (out["d"] <= 0.005)) and np.all(out["S0"] >= 0), j
            assert np.all(np.isfinite(out["th"])) and np.all(np.isfinite(out["ph"])), j
            if model >= 2:
                up = 0.01 if model == 2 else 0.004
                assert np.all((out["dstd"] > 0) & (out["dstd"] <= up)), j
            if model == 3:
                assert np.all((out["R"] > 0) & (out["R"] <= 1)), j
            print(f"job {j} {MIX[j]}: mean f1 {out['mean_f'][0].mean():.3f}, mean S0 {out['mean_S0'].mean():.1f}")
    finally:
        q.close()
