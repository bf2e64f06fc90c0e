"""The learned start on the GPU: the BF16 MFMA network against its CPU
restatement (oracle/mlp_oracle.py) -- outputs and gradients --, training
reducing the loss, the surrogate-initialised run with its plain fallback, and
what the trained network gains over the plain path at the same short burn-in."""
import numpy as np
import pytest

from fdt_amd import synth
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


def rand_targets(g, n):
    t = np.zeros((n, M.SG_OUT), np.float32)
    t[:, 0:3] = g.uniform(0, 0.5, (n, 3))
    for k in range(3):
        x = g.standard_normal((n, 3))
        t[:, 3 + 3 * k:6 + 3 * k] = x / np.linalg.norm(x, axis=1, keepdims=True)
    t[:, 12:14] = g.uniform(0.5, 2.0, (n, 2))
    t[:, 14:] = g.normal(-3, 1, (n, 14))
    t[:, 17] = 1e30
    return t


def truth_targets(v):
    """targets from the synthetic truth (fibres sorted by f): for optimiser checks"""
    n = v["data"].shape[0]
    t = np.full((n, M.SG_OUT), 1e30, np.float32)
    order = np.argsort(-v["f"], axis=1)
    f = np.take_along_axis(v["f"], order, 1)
    dirs = np.take_along_axis(v["dirs"], order[:, :, None], 1)
    t[:, 0:3] = f
    t[:, 3:12] = dirs.reshape(n, 9)
    t[:, 12] = 1000 * v["d"]
    t[:, 13] = 1000 * v["dstd"]
    return t


@pytest.mark.parametrize("name,n", [("c0", 300), ("c1", 129)])
def test_forward_matches_oracle(fdt, name, n):
    v = synth.voxels(name, n)
    s = fdt.Surrogate(v["bvals"].size, seed=3)
    try:
        out = s.forward(v["data"], v["bvals"])
        want, _ = M.forward(M.normalise(v["data"], v["bvals"]), [(W.astype(np.float64), b.astype(np.float64))
                                                                   for W, b in s.layers()])
        print(f"{name}: forward relative error {rel(out, want):.2e}")
        assert rel(out, want) < 1e-2
    finally:
        s.close()


def test_gradients_match_oracle(fdt):
    v = synth.voxels("c1", 257)
    g = np.random.default_rng(7)
    t = rand_targets(g, 257)
    s = fdt.Surrogate(v["bvals"].size, seed=5)
    try:
        layers = [(W.astype(np.float64), b.astype(np.float64)) for W, b in s.layers()]
        s.set_training_data(v["data"], v["bvals"], t)
        gg, loss = s.gradients(0, 257)
        go, loss_o = M.backward(M.normalise(v["data"], v["bvals"]), layers, t, emulate_bf16=True)
        print(f"loss {loss:.6f} oracle {loss_o:.6f}")
        assert loss == pytest.approx(loss_o, rel=1e-2)
        gl = s.layers(gg)
        for l in range(4):
            eW, eb = rel(gl[l][0], go[l][0]), rel(gl[l][1], go[l][1])
            print(f"layer {l}: dW rel err {eW:.2e}, db rel err {eb:.2e}")
            assert eW < 3e-2 and eb < 3e-2
    finally:
        s.close()


def test_training_reduces_loss(fdt):
    from fdt_amd import surrogate as SG
    v = synth.voxels("c1", 8192, sub=3)
    s = fdt.Surrogate(v["bvals"].size, seed=1)
    try:
        losses = SG.train(s, v["data"], v["bvals"], truth_targets(v), epochs=4, batch=256, lr=1e-3, lr_decay=0.7)
        print("losses", [f"{x:.4f}" for x in losses])
        assert losses[-1] < 0.5 * losses[0]
        w = s.weights()
        assert np.all(np.isfinite(w))
    finally:
        s.close()


def test_surrogate_run_untrained_falls_back(fdt):
    v = synth.voxels("c1", 64)
    s = fdt.Surrogate(v["bvals"].size, seed=2)
    x = fdt.Xfibres(nfibres=3, model=2, nburn=200, njumps=200, sampleevery=25)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        nfb = x.run_surrogate(s, nburn_short=50)
        out = x.outputs()
        print(f"fallback voxels: {nfb}/64")
        assert 0 <= nfb <= 64
        f = out["f"]
        assert np.all(np.isfinite(f)) and np.all((f > 0) & (f < 1))
        np.testing.assert_allclose(np.linalg.norm(out["dyads"], axis=1), 1.0, rtol=1e-5)
    finally:
        x.close()
        s.close()


def test_learned_start_gains_over_plain_at_short_burnin(fdt):
    """Trained on the engine's own labels (65,536 voxels), evaluated on other
    voxels: median dyad error vs the truth with burn-in 200 from the learned
    start must not be worse than the plain LM start with the same burn-in, and
    within 0.5 deg of the plain path with the full burn-in of 1000."""
    from fdt_amd import surrogate as SG
    run = dict(nfibres=3, model=2, njumps=1250, sampleevery=25)
    tr = synth.voxels("c1", 65536, sub=1, with_truth=False)
    t = SG.label(tr["bvals"], tr["bvecs"], tr["data"], nburn=1000, **run)
    s = fdt.Surrogate(tr["bvals"].size, seed=11)
    ev = synth.voxels("c1", 4096, sub=2)
    try:
        losses = SG.train(s, tr["data"], tr["bvals"], t, epochs=8, batch=512, lr=1e-3, lr_decay=0.75)
        print("losses", [f"{x:.3f}" for x in losses])

        def errs(out):
            e = SG.dyad_errors(out["dyads"], out["mean_f"], ev["dirs"], ev["f"])
            return np.nanmedian(e), np.mean(np.isnan(e))

        plain_short = fdt.run_xfibres(ev["bvals"], ev["bvecs"], ev["data"], nburn=200, **run)
        plain_full = fdt.run_xfibres(ev["bvals"], ev["bvecs"], ev["data"], nburn=1000, **run)
        x = fdt.Xfibres(nburn=1000, **run)
        try:
            x.set_acquisition(ev["bvals"], ev["bvecs"])
            x.load_data(ev["data"])
            nfb = x.run_surrogate(s, nburn_short=200)
            learned = x.outputs()
        finally:
            x.close()
        ms, ns_ = errs(plain_short)
        mf, nf = errs(plain_full)
        ml, nl = errs(learned)
        print(f"median dyad error: plain bi=200 {ms:.2f} deg (missed {ns_:.3f}), plain bi=1000 {mf:.2f} "
              f"(missed {nf:.3f}), learned start bi=200 {ml:.2f} (missed {nl:.3f}); fallback {nfb}/4096")
        assert ml <= ms + 0.1
        assert ml <= mf + 0.5
    finally:
        s.close()


def test_learned_start_with_f0(fdt):
    """--f0 from the committed multi-shell weights (trained without an f0
    target: the chain starts f0 at 0.001 with the default width): the learned
    path runs every option of the plain one, its fallback fits f0 in the LM,
    and its dyads stay within 0.3 deg (median) of the plain path with f0."""
    import os
    from fdt_amd import surrogate as SG
    ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    wz = np.load(os.path.join(ROOT, "fdt_amd", "weights", "surrogate_c1_model2_nf3.npz"))
    ev = synth.voxels("c1", 4096, sub=3)
    run = dict(nfibres=3, model=2, njumps=1250, sampleevery=25, f0=1, ardf0=1)
    s = fdt.Surrogate(ev["bvals"].size, seed=0)
    try:
        s.load(wz["weights"])
        plain = fdt.run_xfibres(ev["bvals"], ev["bvecs"], ev["data"], nburn=1000, **run)
        x = fdt.Xfibres(nburn=1000, **run)
        try:
            x.set_acquisition(ev["bvals"], ev["bvecs"])
            x.load_data(ev["data"])
            nfb = x.run_surrogate(s, nburn_short=200)
            learned = x.outputs()
        finally:
            x.close()
        assert np.all(np.isfinite(learned["mean_f0"])) and np.all((learned["mean_f0"] >= 0) & (learned["mean_f0"] < 1))
        e_p = SG.dyad_errors(plain["dyads"], plain["mean_f"], ev["dirs"], ev["f"])
        e_l = SG.dyad_errors(learned["dyads"], learned["mean_f"], ev["dirs"], ev["f"])
        print(f"--f0: median dyad error plain {np.nanmedian(e_p):.3f}, learned {np.nanmedian(e_l):.3f} deg; "
              f"mean f0 plain {plain['mean_f0'].mean():.4f}, learned {learned['mean_f0'].mean():.4f}; "
              f"fallback {nfb}/4096")
        assert np.nanmedian(e_l) <= np.nanmedian(e_p) + 0.3
        assert abs(learned["mean_f0"].mean() - plain["mean_f0"].mean()) < 0.02
    finally:
        s.close()
