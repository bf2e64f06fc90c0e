"""Training and evaluation driver of the learned start (fdt_surrogate_* C ABI).

Labels come from this engine's own plain path (LM + MCMC posterior means and
the final adapted proposal widths) on seeded synthetic voxels of the target
acquisition; the network is trained on the GPU by the library's BF16 MFMA
forward/backward kernels and Adam.
"""
from __future__ import annotations

import numpy as np

from . import SG_OUT, SG_WIDTH_SLOTS, Surrogate, Xfibres


def targets_from_outputs(out, widths, model):
    """[n][28] targets from finished plain-path outputs: f_1..3 (fibres sorted
    by mean f), posterior mean dyads, 1000 d, 1000 d_std, log final widths of
    the predicted slots (unused slots marked 1e30)."""
    K = out["mean_f"].shape[0]
    n = out["mean_f"].shape[1]
    t = np.zeros((n, SG_OUT), np.float32)
    for k in range(min(K, 3)):
        t[:, k] = out["mean_f"][k]
        t[:, 3 + 3 * k:6 + 3 * k] = out["dyads"][k].T
    t[:, 12] = 1000.0 * out["mean_d"]
    t[:, 13] = 1000.0 * out["mean_dstd"] if "mean_dstd" in out else 0.0
    # widths of parameters the data do not constrain (angles of an unsupported
    # fibre) grow without bound during adaptation: clip the log targets
    lw = np.clip(np.log(np.maximum(widths[:, SG_WIDTH_SLOTS], 1e-30)), -12.0, 2.0)
    used = np.ones(len(SG_WIDTH_SLOTS), bool)
    used[0] = False                  # f0 (not sampled without --f0)
    used[3] = model == 3             # R
    if model < 2:
        used[2] = False              # d_std
    for k in range(K, 3):
        used[4 + 3 * k + 1:4 + 3 * k + 4] = False
    t[:, 14:] = np.where(used[None, :], lw, 1e30)
    return t


def label(bvals, bvecs, data, seed=8665904, vox_offset=0, **opts):
    """plain-path labels for the voxels `data` [n][N]"""
    x = Xfibres(seed=seed, **opts)
    try:
        x.set_acquisition(bvals, bvecs)
        x.load_data(data)
        x.run(vox_offset)
        x.sync()
        out = x.outputs(raw_scalars=False)
        w = x.widths()
    finally:
        x.close()
    return targets_from_outputs(out, w, opts.get("model", 1))


def train(s: Surrogate, signals, bvals, targets, epochs=4, batch=1024, lr=1e-3, seed=0, lr_decay=0.5, log=None):
    """Adam over shuffled mini-batches; the learning rate is multiplied by
    lr_decay after every epoch.  Returns the mean loss of each epoch."""
    g = np.random.default_rng(seed)
    n = signals.shape[0]
    losses = []
    for e in range(epochs):
        perm = g.permutation(n)
        s.set_training_data(signals[perm], bvals, targets[perm])
        tot, cnt = 0.0, 0
        for first in range(0, n - batch + 1, batch):
            tot += s.train_step(first, batch, lr)
            cnt += 1
        losses.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {e}: loss {losses[-1]:.4f} lr {lr:.2e}")
        lr *= lr_decay
    return losses


def dyad_errors(dyads, mean_f, truth_dirs, truth_f, fmin=0.05):
    """angle (deg) between each true fibre (f > 0) and the estimated fibre
    (mean f > fmin) that matches it, matched greedily in the order of the true
    fibres by |cos|; NaN where no supported estimate is left.  dyads
    [K][3][n], mean_f [K][n], truth_dirs [n][3][3], truth_f [n][3]."""
    K, _, n = dyads.shape
    est = np.transpose(dyads, (2, 0, 1)).astype(np.float64)          # [n][K][3]
    ok = (mean_f.T > fmin)                                            # [n][K]
    cos = np.abs(np.einsum("nkc,ntc->ntk", est, truth_dirs))          # [n][true][K]
    out = np.full((n, 3), np.nan)
    for kt in range(3):
        c = np.where(ok, cos[:, kt, :], -1.0)
        bi = np.argmax(c, axis=1)
        best = c[np.arange(n), bi]
        has = (best >= 0) & (truth_f[:, kt] > 0)
        out[has, kt] = np.degrees(np.arccos(np.minimum(1.0, best[has])))
        ok[np.arange(n)[has], bi[has]] = False
        out[(truth_f[:, kt] > 0) & ~has, kt] = np.nan
    return out[truth_f > 0]
