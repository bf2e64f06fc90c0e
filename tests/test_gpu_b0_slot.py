"""Acquisitions with at least 16 b = 0 volumes: the chains stage those volumes
first (a stable partition of the acquisition table, per workgroup) and the
moves that leave them unchanged skip their slot.  The synthetic acquisitions
already put their b = 0 volumes first, so the partition is the identity there;
here the 18 b = 0 volumes of the multi-shell acquisition are spread through the
diffusion-weighted ones (relative orders kept).  Both orders stage the same
table, so the interleaved acquisition must give exactly the chains of the
b = 0-first one, Gaussian, Rician and --gradnonlin, and the batched queue must
stage it as the per-volume launch does."""
import numpy as np
import pytest

from fdt_amd import synth

pytestmark = pytest.mark.gpu


def interleave(v):
    """v's acquisition and data with a b = 0 volume at every 16th position from 5
    on (the b = 0 and the diffusion-weighted volumes each keep their order)"""
    b = v["bvals"]
    z, nz = list(np.flatnonzero(b == 0)), list(np.flatnonzero(b != 0))
    order, pos = [], 0
    while z or nz:
        order.append(z.pop(0) if z and (pos % 16 == 5 or not nz) else nz.pop(0))
        pos += 1
    P = np.array(order)
    assert sorted(P) == list(range(b.size)) and P[0] != 0
    return dict(bvals=b[P], bvecs=np.ascontiguousarray(v["bvecs"][:, P]), data=np.ascontiguousarray(v["data"][:, P]))


def chains(fdt, a, params, tau, grad=None, **run):
    x = fdt.Xfibres(**run)
    try:
        x.set_acquisition(a["bvals"], a["bvecs"])
        x.load_data(a["data"])
        if grad is not None:
            x.load_grad_dev(grad)
        x.set_params(params, tau)
        x.mcmc(0)
        e = x.energy()
        x.finish()
        return x.outputs(), e
    finally:
        x.close()


def first_volumes(v, n):
    return dict(bvals=v["bvals"][:n].copy(), bvecs=np.ascontiguousarray(v["bvecs"][:, :n]),
                data=np.ascontiguousarray(v["data"][:, :n]))


# N = 288 fills the 18-volume tile; N = 200 leaves 88 padding slots after the partition
@pytest.mark.parametrize("model,rician,N", [(2, 0, 288), (3, 1, 288), (2, 0, 200)])
def test_interleaved_b0_volumes_give_the_b0_first_chains(fdt, model, rician, N):
    v = first_volumes(synth.voxels("c1", 96), N)
    assert int((v["bvals"] == 0).sum()) == 18 and (v["bvals"][:18] == 0).all()
    w = interleave(v)
    run = dict(nfibres=3, model=model, rician=rician, nburn=100, njumps=250, sampleevery=25)
    x = fdt.Xfibres(**run)
    try:  # one LM start for both orders
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        x.fit()
        params, tau = x.get_params()
    finally:
        x.close()
    tau = tau if rician else None
    want, e_want = chains(fdt, v, params, tau, **run)
    got, e_got = chains(fdt, w, params, tau, **run)
    assert np.isfinite(want["th"]).all() and np.isfinite(e_want).all()
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    np.testing.assert_array_equal(e_got, e_want)


# one wave (4 voxels) or four waves (16 voxels) per workgroup: the staging runs on whole waves
@pytest.mark.parametrize("rows_per_wg", [4, 16])
def test_queue_stages_interleaved_b0_volumes_like_the_per_volume_launch(fdt, rows_per_wg):
    w = interleave(synth.voxels("c2_288", 45))
    kw = dict(nfibres=3, model=2, nburn=40, njumps=100, sampleevery=25)
    q = fdt.Queue()
    try:
        q.enqueue(w["bvals"], w["bvecs"], w["data"], vox_offset=0, **kw)
        q.set_tiles(rows_per_wg=rows_per_wg, job=0, vpl=18)
        q.run()
        assert q.timing()["rows_per_wg"] == rows_per_wg
        got = q.outputs(0)
        want = fdt.run_xfibres(w["bvals"], w["bvecs"], w["data"], vox_offset=0, **kw)
        for k in ("th", "ph", "f", "mean_S0", "mean_d", "dyads"):
            np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    finally:
        q.close()


def test_interleaved_b0_volumes_gradnonlin(fdt):
    """--gradnonlin: the per-voxel b values are zero exactly where the shared
    table's are, and the rows read them in the staged order"""
    v = synth.voxels("c1", 64)
    w = interleave(v)
    grad = synth.grad_dev(np.random.default_rng(5), 64)
    run = dict(nfibres=3, model=2, nburn=100, njumps=250, sampleevery=25)
    x = fdt.Xfibres(**run)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        x.load_grad_dev(grad)
        x.fit()
        params, _ = x.get_params()
    finally:
        x.close()
    want, e_want = chains(fdt, v, params, None, grad=grad, **run)
    got, e_got = chains(fdt, w, params, None, grad=grad, **run)
    assert np.isfinite(want["th"]).all()
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    np.testing.assert_array_equal(e_got, e_want)
