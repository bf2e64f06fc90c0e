"""Posterior prediction (xfibres_pred) on the GPU against its restatement
(oracle/pred_oracle.py, pinned in tests/test_oracle_pred.py): the kernel
through the C ABI, and the xfibres_pred program on a bedpostx_gpu result."""
import os
import subprocess

import numpy as np
import pytest

import nifti_np as NI
from fdt_amd import synth
from oracle import pred_oracle as P
from test_gpu_pipeline import PIPE, make_subject
from test_oracle_pred import rand_voxels

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PRED = os.path.join(ROOT, "fdt_amd", "bin", "xfibres_pred")


@pytest.mark.parametrize("model,use_f0", [(1, False), (1, True), (2, False), (2, True), (3, False), (3, True)])
def test_kernel_matches_oracle(fdt, model, use_f0):
    g = np.random.default_rng(model + 7 * use_f0)
    K, n = 3, 3000
    v = rand_voxels(n, K, g)
    v["dstd"][::7] = 5e-6          # the small-d_std forms
    v["f"][2, ::5] = 0.0           # unsupported third fibre
    bvals, bvecs = synth.acquisition(5, [(1000.0, 40), (2000.0, 40)], g)
    args = (v["S0"], v["d"], v["f"], v["dyads"], model, v["dstd"], v["R"], v["f0"] if use_f0 else None)
    got = fdt.predict(bvals, bvecs, *args)
    want = P.predict(bvals.astype(np.float32).astype(np.float64), bvecs.astype(np.float32).astype(np.float64),
                     *[a.astype(np.float32) if isinstance(a, np.ndarray) else a for a in args])
    np.testing.assert_array_max_ulp(got, want, maxulp=2)
    print(f"model {model} f0 {use_f0}: identical {np.mean(got == want):.4f}")


def test_empty_and_errors(fdt):
    out = fdt.predict(np.zeros(3), np.eye(3), np.zeros(0), np.zeros(0), np.zeros((1, 0)), np.zeros((1, 3, 0)))
    assert out.shape == (3, 0)
    with pytest.raises(RuntimeError):
        fdt.predict(np.zeros(3), np.eye(3), np.ones(2), np.ones(2), np.ones((1, 2)), np.ones((1, 3, 2)), model=2)


def test_xfibres_pred_on_a_bedpostx_result(fdt, tmp_path):
    s, v, (xs, ys, zs) = make_subject(tmp_path)
    r = subprocess.run([PIPE, str(s), "-n", "2", "-b", "200", "-j", "250", "-s", "25", "-model", "1"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    B = tmp_path / "subj.bedpostX"
    out = tmp_path / "pred"
    print(PRED, B, out, flush=True)
    r = subprocess.run([PRED, str(B), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Files for model1 with 2 fibres found" in r.stdout
    pred, hdr = NI.read(str(out) + ".nii.gz")
    N = v["bvals"].size
    assert pred.shape == (9, 8, 4, N)
    m = (xs, ys, zs)
    rd = lambda name: NI.read(B / (name + ".nii.gz"))[0]
    f = np.stack([rd("mean_f1samples")[m], rd("mean_f2samples")[m]])
    dy = np.stack([rd("dyads1")[m].T, rd("dyads2")[m].T])
    bv = np.loadtxt(B / "bvals").ravel()
    bc = np.loadtxt(B / "bvecs")
    bc = (bc / np.where(np.linalg.norm(bc, axis=0) > 0, np.linalg.norm(bc, axis=0), 1)).astype(np.float32)
    want = P.predict(bv.astype(np.float32).astype(np.float64), bc.astype(np.float64), rd("mean_S0samples")[m],
                     rd("mean_dsamples")[m], f, dy, 1)
    np.testing.assert_array_max_ulp(pred[m].T, want, maxulp=2)
    # a posterior-predictive check: the prediction explains the data to the noise level (SNR 20)
    rel = np.sqrt(np.mean((pred[m] - v["data"]) ** 2, axis=1)) / rd("mean_S0samples")[m]
    assert np.median(rel) < 0.08, np.median(rel)
    # without mean_S0samples: S0 from the b <= 50 volumes of the data file
    (B / "mean_S0samples.nii.gz").unlink()
    r = subprocess.run([PRED, str(B), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "No mean_S0samples or data file exists!" in r.stderr
    r = subprocess.run([PRED, str(B), str(s / "data"), str(tmp_path / "pred2.nii.gz")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "getting the average of b=0's" in r.stdout
    pred2, _ = NI.read(tmp_path / "pred2.nii.gz")
    data, _ = NI.read(s / "data.nii.gz")
    b0 = (bv >= 0) & (bv <= 50)
    S0 = np.zeros(data.shape[:3], np.float32)
    for l in np.nonzero(b0)[0]:
        S0 = (S0 + data[..., l]).astype(np.float32)
    S0 = (S0 / np.float32(b0.sum())).astype(np.float32)
    want2 = P.predict(bv.astype(np.float32).astype(np.float64), bc.astype(np.float64), S0[m], rd("mean_dsamples")[m], f,
                      dy, 1)
    np.testing.assert_array_max_ulp(pred2[m].T, want2, maxulp=2)
