"""Pins of the oracle's --gradnonlin path: correct_bvals_bvecs
(CUDA/xfibres_gpu.cu:151-170; xfibres.cc:797-817) against what the geometry
fixes, and the per-voxel acquisition reaching the fit and the chain
(CUDA/xfibres_gpu.cu:403-436; xfibres.cc:909-919)."""
import numpy as np
import pytest

from fdt_amd import synth


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def acq():
    bvals, bvecs = synth.acquisition(3, [(1000.0, 20), (2500.0, 10)], synth.rng("test", 77))
    return f32(bvals), f32(bvecs)


def grad_of(M):
    """the 9 grad_dev values of the deviation L = M - I (L(i, j) = g[3j + i])"""
    L = np.asarray(M, dtype=np.float64) - np.eye(3)
    return L.T.reshape(9).copy()


def test_zero_deviation_is_identity_to_float_rounding(orc, acq):
    # |v| of a float-rounded unit vector rounds to 1 or to 1 - 2^-24
    bv, bc = acq
    b2, c2 = orc.correct_acq(bv, bc, np.zeros(9))
    np.testing.assert_allclose(b2, bv, rtol=2.5e-7)
    np.testing.assert_allclose(c2, bc, atol=1.5e-7)


@pytest.mark.parametrize("s", [0.05, -0.08])
def test_isotropic_scaling_scales_b(orc, acq, s):
    bv, bc = acq
    b2, c2 = orc.correct_acq(bv, bc, grad_of((1 + s) * np.eye(3)))
    dw = bv > 0
    np.testing.assert_allclose(b2[dw], bv[dw] * (1 + s) ** 2, rtol=3e-7)
    np.testing.assert_allclose(c2[:, dw], bc[:, dw], atol=2e-7)
    np.testing.assert_array_equal(b2[~dw], bv[~dw])


def test_rotation_rotates_directions_and_keeps_b(orc, acq):
    bv, bc = acq
    a = np.radians(20.0)
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])
    b2, c2 = orc.correct_acq(bv, bc, grad_of(R))
    dw = bv > 0
    np.testing.assert_allclose(b2[dw], bv[dw], rtol=3e-7)
    np.testing.assert_allclose(c2[:, dw], (R @ bc)[:, dw], atol=3e-7)


def test_index_convention_shear(orc):
    # grad[1] is L(2,1): the y component of the deviated x gradient
    bv = np.array([0.0, 1000.0, 1000.0])
    bc = np.array([[1.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]])
    g = np.zeros(9)
    g[1] = 0.25
    b2, c2 = orc.correct_acq(bv, bc, g)
    np.testing.assert_allclose(c2[:, 1], np.array([1.0, 0.25, 0.0]) / np.hypot(1.0, 0.25), atol=1e-7)
    assert b2[1] == pytest.approx(1000.0 * (1 + 0.25 ** 2), rel=1e-7)
    np.testing.assert_array_equal(c2[:, 2], bc[:, 2])           # y gradient: no x-column term
    assert b2[2] == 1000.0
    np.testing.assert_array_equal(c2[:, 0], bc[:, 0])           # b = 0 volume untouched
    assert b2[0] == 0.0


def test_b0_volumes_untouched(orc, acq):
    bv, bc = acq
    g = synth.grad_dev(np.random.default_rng(5), 1, 0.2)[0]
    b2, c2 = orc.correct_acq(bv, bc, g)
    b0 = bv == 0
    np.testing.assert_array_equal(b2[b0], bv[b0])
    np.testing.assert_array_equal(c2[:, b0], bc[:, b0])
    assert np.all(b2[~b0] != bv[~b0])


def test_matches_vectorised_generator(orc, acq):
    bv, bc = acq
    grad = synth.grad_dev(np.random.default_rng(6), 5, 0.05)
    B, Cc = synth.corrected_acquisitions(bv, bc, grad)
    for v in range(5):
        b2, c2 = orc.correct_acq(bv, bc, grad[v].astype(np.float64))
        np.testing.assert_allclose(b2, B[v], rtol=3e-7)
        np.testing.assert_allclose(c2, Cc[v], atol=3e-7)


def test_each_voxel_is_fitted_and_sampled_with_its_corrected_acquisition(orc):
    # the --gradnonlin result of voxel i is the plain result of voxel i alone
    # on the acquisition correct_acq(grad_i): same fit, same chain, bit for bit
    n = 6
    v = synth.gradnonlin_voxels("c0", n, sd=0.05)
    data = v["data"].astype(np.float64)
    bv, bc = f32(v["bvals"]), f32(v["bvecs"])
    o = orc.opts(nfibres=2, model=1, njumps=100, nburn=50, sampleevery=25)
    pg, tg, _ = orc.fit(o, bv, bc, data, grad=v["grad"])
    sg = orc.mcmc(o, bv, bc, data, pg, vox_offset=40, grad=v["grad"])
    for i in range(n):
        b2, c2 = orc.correct_acq(bv, bc, v["grad"][i].astype(np.float64))
        pi, _, _ = orc.fit(o, b2, c2, data[i:i + 1])
        np.testing.assert_array_equal(pi[0], pg[i])
        si = orc.mcmc(o, b2, c2, data[i:i + 1], pi, vox_offset=40 + i)
        for k in ("th", "ph", "f"):
            np.testing.assert_array_equal(si[k][..., 0], sg[k][..., i])
        for k in ("S0", "d"):
            np.testing.assert_array_equal(si[k][:, 0], sg[k][:, i])
    st = orc.start(orc.K_SINGLE_C, 2, False, bv, bc, data, grad=v["grad"])
    b2, c2 = orc.correct_acq(bv, bc, v["grad"][3].astype(np.float64))
    np.testing.assert_array_equal(orc.start(orc.K_SINGLE_C, 2, False, b2, c2, data[3:4])[0], st[3])


@pytest.mark.parametrize("model,K,name", [(1, 2, "c0"), (2, 3, "c1")])
def test_fit_uses_the_voxels_own_acquisition(orc, model, K, name):
    # noise-free data acquired through per-voxel deviations: the fit with the
    # deviations recovers the truth, the fit without them does not
    n = 32
    v = synth.gradnonlin_voxels(name, n, sd=0.04, clean_only=True)
    data = v["clean"].astype(np.float64)
    bv, bc = f32(v["bvals"]), f32(v["bvecs"])
    o = orc.opts(nfibres=K, model=model)
    pg, _, _ = orc.fit(o, bv, bc, data, grad=v["grad"])
    pn, _, _ = orc.fit(o, bv, bc, data)
    eg = np.abs(pg[:, 1] / v["d"] - 1)
    en = np.abs(pn[:, 1] / v["d"] - 1)
    if model == 1:
        assert np.mean(eg < 1e-3) >= 0.9, np.sort(eg)[-5:]
        assert np.median(en) > 10 * max(np.median(eg), 1e-6), (np.median(en), np.median(eg))
    else:
        # three sticks on 1-3 true fibres: PVM_multi lands exactly on the truth
        # for a part of the voxels only, but never without the deviations
        gs = np.abs(pg[:, 0] / v["S0"] - 1)
        ns = np.abs(pn[:, 0] / v["S0"] - 1)
        assert np.mean(gs < 1e-5) >= 0.2 and np.mean(ns < 1e-5) == 0.0
