"""The measured workload and its algorithmic flop model."""
import numpy as np

import bench
from fdt_amd import synth


def test_flop_model_matches_per_sweep_figures():
    # per-voxel, per-sweep figures of the workloads: model/K/N -> flops
    assert bench.flops_per_jump(1, 2, 69) == 9108
    assert bench.flops_per_jump(2, 3, 288) == 82944
    assert bench.flops_per_jump(3, 3, 288, rician=True) == 188352
    assert [bench.flops_per_jump(m, k, n) for m, k, n in [(1, 1, 33), (2, 2, 69), (2, 3, 145), (3, 3, 288)]] == \
        [2244, 13179, 41760, 79776]
    assert bench.flops_per_jump(2, 3, 288) * 2250 == 186_624_000


def test_c1_mask_and_acquisition():
    assert synth.C1_NVOX == 250968
    v = synth.voxels("c1", 8)
    assert v["bvals"].size == 288 and (v["bvals"] == 0).sum() == 18
    for b in (1000.0, 2000.0, 3000.0):
        assert (v["bvals"] == b).sum() == 90
    np.testing.assert_allclose(np.linalg.norm(v["bvecs"], axis=0), 1.0, rtol=1e-12)


def test_synthetic_is_seeded_and_block_stable():
    a = synth.voxels("c1", 40)
    b = synth.voxels("c1", 40)
    np.testing.assert_array_equal(a["data"], b["data"])
    big = synth.voxels("c0", synth.CHUNK + 5, with_truth=False)
    one = synth.voxels("c0", synth.CHUNK, with_truth=False)
    np.testing.assert_array_equal(big["data"][:synth.CHUNK], one["data"])
    c = synth.voxels("c0", 10)
    assert c["data"].shape == (10, 69)
    assert np.all(c["f"].sum(1) < 1) and np.all(c["f"][:, 0] >= 0.3)


def test_bench_refuses_more_than_one_gpu():
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU per process" in r.stderr


def test_bench_helpers():
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    # the flop model of the default workload: 82,944 per voxel and sweep (model 2, K=3, N=288)
    assert bench.flops_per_jump(2, 3, 288) == 82944
    # the learned start as the learned-start workload specifies it (burn-in 200), with the burn-in-50 variant
    assert bench.SG_DEFAULTS["nburn_short"] == 200 and bench.SG_SHORT["nburn_short"] == 50
    assert bench.SG_DEFAULTS_C4["nburn_short"] == 200
    # LM flop model at the multi-shell size: 288 * (12*13 + 24 + 120) per loop trip
    assert bench.lm_flops_per_iter(288, 12) == 288 * 300
    assert bench.LM_DEFAULT in bench.LM_MODES
