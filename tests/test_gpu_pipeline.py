"""The in-process bedpostx_gpu pipeline end to end on the GPU: a NIfTI subject
directory in, <subject>.bedpostX out.  The merged samples must be the
library's samples of the masked voxels (bit for bit), and the post-processed
images (-Tmean, dyads, dispersion, maskdyads, modf, mean_fsumsamples) must
match the post-processing oracle computed from those samples."""
import os
import subprocess

import numpy as np
import pytest

import nifti_np as NI
from fdt_amd import synth
from oracle import postproc_oracle as PP

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PIPE = os.path.join(ROOT, "fdt_amd", "bin", "bedpostx_gpu")
SHAPE = (9, 8, 4)


def make_subject(root, grad=False):
    """a 9x8x4 grid, an ellipsoidal mask, c0 voxels (N=69) inside it"""
    x, y, z = np.meshgrid(*[np.arange(n) for n in SHAPE], indexing="ij")
    mask = (((x - 4) / 4.2) ** 2 + ((y - 3.5) / 3.7) ** 2 + ((z - 1.5) / 2.2) ** 2 <= 1.0)
    nvox = int(mask.sum())
    v = synth.gradnonlin_voxels("c0", nvox, sd=0.03) if grad else synth.voxels("c0", nvox)
    N = v["bvals"].size
    # mask voxels in FSL order (x fastest, then y, then z)
    xs, ys, zs = np.nonzero(mask.transpose(2, 1, 0))[::-1]
    data = np.zeros(SHAPE + (N,), np.float32)
    data[xs, ys, zs, :] = v["data"]
    s = root / "subj"
    s.mkdir()
    NI.write(s / "data.nii.gz", data)
    NI.write(s / "nodif_brain_mask.nii.gz", mask.astype(np.uint8))
    NI.write(s / "nodif.nii.gz", data[..., 0])
    np.savetxt(s / "bvals", v["bvals"][None, :])
    np.savetxt(s / "bvecs", v["bvecs"])
    if grad:
        g = np.zeros(SHAPE + (9,), np.float32)
        g[xs, ys, zs, :] = v["grad"]
        NI.write(s / "grad_dev.nii.gz", g)
    return s, v, (xs, ys, zs)


def library_run(fdt, v, grad):
    return fdt.run_xfibres(v["bvals"], v["bvecs"], v["data"], nfibres=2, model=1, nburn=50, njumps=100,
                           sampleevery=25, cnonlin=1, grad=v["grad"] if grad else None)


@pytest.mark.parametrize("grad", [False, True])
def test_pipeline_end_to_end(fdt, tmp_path, grad):
    s, v, (xs, ys, zs) = make_subject(tmp_path, grad)
    cmd = [PIPE, str(s), "-n", "2", "-b", "50", "-j", "100", "-s", "25", "-model", "1", "-NJOBS", "4"] + \
        (["-g"] if grad else [])
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Done" in r.stdout
    B = tmp_path / "subj.bedpostX"
    ref = library_run(fdt, v, grad)
    ns = 4
    rd = lambda name: NI.read(B / (name + ".nii.gz"))
    for k in (1, 2):
        for q, key in (("th", "th"), ("ph", "ph"), ("f", "f")):
            img, hdr = rd(f"merged_{q}{k}samples")
            assert img.shape == SHAPE + (ns,) and hdr["datatype"] == 16
            np.testing.assert_array_equal(img[xs, ys, zs, :], ref[key][k - 1].T)
            outside = np.ones(SHAPE, bool)
            outside[xs, ys, zs] = False
            assert not img[outside].any()
            mean, _ = rd(f"mean_{q}{k}samples")
            np.testing.assert_allclose(mean[xs, ys, zs], PP.tmean(ref[key][k - 1].T), rtol=1e-6, atol=1e-7)
        th, _ = rd(f"merged_th{k}samples")
        ph, _ = rd(f"merged_ph{k}samples")
        want_d, want_disp = PP.dyadic_vectors(th[xs, ys, zs, :], ph[xs, ys, zs, :])
        dy, hdr = rd(f"dyads{k}")
        assert hdr["cal"] == (-1.0, 1.0)
        got = dy[xs, ys, zs, :]
        np.testing.assert_allclose(np.abs((got * want_d).sum(1)), 1.0, atol=2e-5)
        disp, _ = rd(f"dyads{k}_dispersion")
        np.testing.assert_allclose(disp[xs, ys, zs], want_disp, atol=2e-5)
    f1, _ = rd("mean_f1samples")
    f2, _ = rd("mean_f2samples")
    dy2, _ = rd("dyads2")
    thr, _ = rd("dyads2_thr0.05")
    modf, _ = rd("dyads2_thr0.05_modf2")
    m = (xs, ys, zs)
    want_thr = PP.maskdyads(dy2[m], f2[m])
    np.testing.assert_array_equal(thr[m], want_thr)
    np.testing.assert_array_equal(modf[m], PP.modf(want_thr, f2[m], f1[m]))
    fs, _ = rd("mean_fsumsamples")
    np.testing.assert_array_equal(fs[m], PP.fsum([f1[m], f2[m]]))
    for name in ("mean_dsamples", "mean_S0samples"):
        img, hdr = rd(name)
        assert img.shape == SHAPE
        assert hdr["cal"][1] == img.max()
    np.testing.assert_array_equal(rd("mean_S0samples")[0][m], ref["mean_S0"])
    assert (B / "xfms" / "eye.mat").read_text().split() == "1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1".split()
    assert (B / "bvals").exists() and (B / "bvecs").exists() and (B / "nodif_brain_mask.nii.gz").exists()
    nb, _ = NI.read(B / "nodif_brain.nii.gz")
    np.testing.assert_array_equal(nb[m], v["data"][:, 0])
    # a second run refuses to overwrite
    r2 = subprocess.run(cmd, capture_output=True, text=True, timeout=60)
    assert r2.returncode == 1 and "has already been processed" in r2.stderr


def test_pipeline_cpu_initialisation(fdt, tmp_path):
    """--cpuinit: the chains start as the CPU xfibres starts them (neighbour initialisation over the mask
    grid, tensor fallback); the merged samples equal the library's neighbour-initialised run on that grid"""
    s, v, (xs, ys, zs) = make_subject(tmp_path)
    cmd = [PIPE, str(s), "-n", "2", "-b", "50", "-j", "100", "-s", "25", "-model", "1", "--cpuinit"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    x = fdt.Xfibres(nfibres=2, model=1, nburn=50, njumps=100, sampleevery=25, cpu_init=1, seed=8665904)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        x.set_grid(SHAPE, xs + SHAPE[0] * (ys + SHAPE[1] * zs))
        x.run()
        x.sync()
        ref = x.outputs()
        st = x.spatial_state()
    finally:
        x.close()
    assert (st["been"] >= 2).any()
    B = tmp_path / "subj.bedpostX"
    for k in (1, 2):
        for q in ("th", "ph", "f"):
            img, _ = NI.read(B / f"merged_{q}{k}samples.nii.gz")
            np.testing.assert_array_equal(img[xs, ys, zs, :], ref[q][k - 1].T)
