"""The batched queue (fdt_enqueue_batch / fdt_run_queue): volumes with
different acquisitions, models and fibre counts run in one persistent chain
launch and must give what each volume gives on its own (same tiles, same
seeds and voxel offsets); graph replays must repeat the first run; the
autotuner must leave a working configuration."""
import numpy as np
import pytest

from fdt_amd import synth

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
    q, vols = build_queue(fdt, 48)
    try:
        ms = q.autotune(jumps=20, sample_vox=48)
        t = q.timing()
        print(f"autotune: best {ms:.3f} ms, rows/wg {t['rows_per_wg']}, tiles {t['vpl']}")
        assert ms > 0 and t["rows_per_wg"] in (4, 8, 16)
        for j, (name, model, K) in enumerate(MIX):
            assert 16 * t["vpl"][j] >= vols[j][0]["bvals"].size
        q.run()
        for j, (name, model, K) in enumerate(MIX):
            out = q.outputs(j)
            f = out["f"]
            assert np.all(np.isfinite(f)) and np.all((f > 0) & (f < 1))
            np.testing.assert_allclose(np.linalg.norm(out["dyads"], axis=1), 1.0, rtol=1e-5)
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
