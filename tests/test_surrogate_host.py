"""Host-side pieces of the learned start and of the high-resolution workload:
the training targets built from finished outputs, the dyad-error matching, and
the synthetic model-3 volume (SNR field, narrow crossings)."""
import numpy as np
import pytest

from fdt_amd import synth
from fdt_amd import surrogate as SG
import fdt_amd as F


def fake_outputs(n, K, model, g):
    f = np.sort(g.uniform(0.05, 0.3, (K, n)), axis=0)[::-1]
    d = g.standard_normal((K, 3, n))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out = {"mean_f": f.astype(np.float32), "dyads": d.astype(np.float32),
           "mean_d": g.uniform(1e-3, 2e-3, n).astype(np.float32)}
    if model >= 2:
        out["mean_dstd"] = g.uniform(1e-4, 5e-4, n).astype(np.float32)
    return out


@pytest.mark.parametrize("model,K", [(1, 2), (2, 3), (3, 3)])
def test_targets_layout(model, K):
    g = np.random.default_rng(model)
    n = 7
    out = fake_outputs(n, K, model, g)
    w = np.exp(g.uniform(-8, 1, (n, 16))).astype(np.float32)
    w[:, 6] = 1e10  # an unconstrained angle: clipped
    t = SG.targets_from_outputs(out, w, model)
    assert t.shape == (n, F.SG_OUT)
    np.testing.assert_allclose(t[:, :K], out["mean_f"].T)
    np.testing.assert_allclose(t[:, 3:6], out["dyads"][0].T)
    np.testing.assert_allclose(t[:, 12], 1000 * out["mean_d"], rtol=1e-6)
    lw = t[:, 14:]
    used = lw < 1e29
    assert not used[:, 0].any()                      # f0 slot
    assert used[:, 3].all() == (model == 3)          # R slot
    assert used[:, 2].all() == (model >= 2)          # d_std slot
    assert used[:, 1].all() and used[:, 4].all()     # d, S0
    assert lw[:, 5].max() <= 2.0                     # clipped theta_1 width
    for k in range(K, 3):
        assert not used[:, 5 + 3 * k:8 + 3 * k].any()
    np.testing.assert_allclose(lw[:, 1], np.log(w[:, 2]), rtol=1e-5)


def test_dyad_errors_matching():
    g = np.random.default_rng(3)
    n = 40
    td = g.standard_normal((n, 3, 3))
    td /= np.linalg.norm(td, axis=2, keepdims=True)
    tf = np.where(g.uniform(size=(n, 3)) < 0.7, 0.3, 0.0)
    est = -td[:, [2, 0, 1], :]                      # permuted and sign-flipped, exact
    mf = np.full((3, n), 0.3)
    e = SG.dyad_errors(np.transpose(est, (1, 2, 0)), mf, td, tf)
    assert e.shape == ((tf > 0).sum(),)
    assert np.nanmax(e) < 1e-4
    mf[:, :] = 0.0                                  # nothing supported: every true fibre unmatched
    e = SG.dyad_errors(np.transpose(est, (1, 2, 0)), mf, td, tf)
    assert np.all(np.isnan(e))
    # a known angle
    est2 = td.copy()
    ang = np.radians(10.0)
    v = td[:, 0]
    w = np.cross(v, [0.0, 0.0, 1.0])
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    est2[:, 0] = np.cos(ang) * v + np.sin(ang) * w
    tf1 = np.zeros((n, 3))
    tf1[:, 0] = 0.5
    e = SG.dyad_errors(np.transpose(est2, (1, 2, 0)), np.full((3, n), 0.3), td, tf1)
    np.testing.assert_allclose(e, 10.0, atol=1e-6)


def test_c4_volume_properties():
    assert synth.C4_NVOX == 904550
    v = synth.voxels("c4", 3000)
    assert v["model"] == 3 and v["data"].shape == (3000, 288)
    assert np.all((v["snr"] >= 10) & (v["snr"] <= 30))
    two = v["nfib_true"] >= 2
    ang = np.degrees(np.arccos(np.abs((v["dirs"][:, 0] * v["dirs"][:, 1]).sum(1))))
    frac = np.mean(ang[two] < 45)
    assert 0.10 < frac < 0.20
    assert np.all(v["dstd"] <= 0.004 + 1e-12)
    # the SNR field is smooth: grid neighbours differ little
    pos = synth.c4_positions()[::97]
    s = synth.c4_snr(pos)
    for ax in range(3):
        step = np.zeros(3)
        step[ax] = 1.0
        assert np.max(np.abs(synth.c4_snr(pos + step) - s)) < 0.2
    # model-3 sticks keep the trace: over three orthogonal gradient directions
    # the mean log attenuation of a pure stick is -b d, whatever the direction
    b = np.full(3, 1000.0)
    u = np.linalg.qr(np.random.default_rng(1).standard_normal((3, 3)))[0]
    sig = synth.clean_signal(3, b, u, np.array([2.0]), np.array([1.2e-3]), np.array([3e-4]),
                             np.array([[1.0]]), np.array([[[0.3, -0.4, np.sqrt(0.75)]]]))
    assert np.mean(np.log(sig[0] / 2.0)) == pytest.approx(-1.2, rel=1e-12)
    assert np.ptp(sig[0]) > 0.01   # but anisotropic
