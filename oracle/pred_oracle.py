"""CPU restatement of xfibres_pred (test infrastructure only).

The posterior-predictive signal of xfibres_pred.cc:153-190 from the posterior
means of the masked voxels, with the reference's promotions: bvals / bvecs
are double (newmat), the images float; the projection `angp` is a double sum
rounded to float, exponents in double, every `output += term` rounds to
float, the model-3 factor (1-R) angp^2 + R is float, and the result is scaled
by S0 in float.  Only tests/ import this module.
"""
import numpy as np

f32 = np.float32
f64 = np.float64


def predict(bvals, bvecs, S0, d, f, dyads, model=1, dstd=None, R=None, f0=None):
    """bvals [N], bvecs [3][N] (unit), per-voxel S0, d, dstd, R, f0 [nvox],
    f [K][nvox], dyads [K][3][nvox] -> [N][nvox] float32"""
    b = np.asarray(bvals, f64)[:, None]
    g = np.asarray(bvecs, f64)
    S0, d = np.asarray(S0, f32), np.asarray(d, f32)
    f, dyads = np.asarray(f, f32), np.asarray(dyads, f32)
    K, nvox = f.shape
    small = np.zeros(nvox, bool) if model == 1 else (np.asarray(dstd, f32) <= f32(1e-5))
    stick1 = np.ones(nvox, bool) if model == 1 else (small if model == 2 else np.zeros(nvox, bool))
    ball1 = np.ones(nvox, bool) if model == 1 else small
    sig2 = ((np.asarray(dstd, f32) * np.asarray(dstd, f32)).astype(f32) if model >= 2 else np.zeros(nvox, f32))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        dalpha = ((d * d).astype(f32) / sig2).astype(f32) if model >= 2 else np.zeros(nvox, f32)
    Rv = np.asarray(R, f32) if model == 3 else np.zeros(nvox, f32)
    out = np.zeros((len(bvals), nvox), f32)
    sumf = np.zeros(nvox, f32)
    dd, s2, da_ = d.astype(f64), sig2.astype(f64), dalpha.astype(f64)
    for k in range(K):
        sumf = (sumf + f[k]).astype(f32)
        angp = ((dyads[k, 0].astype(f64)[None, :] * g[0][:, None] + dyads[k, 1].astype(f64)[None, :] * g[1][:, None])
                + dyads[k, 2].astype(f64)[None, :] * g[2][:, None]).astype(f32)
        a64 = angp.astype(f64)
        e1 = np.exp(((-b * dd) * a64) * a64)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            e2 = np.exp(da_ * np.log(dd / (dd + ((b * a64) * a64) * s2)))
        t = ((((f32(1) - Rv) * angp).astype(f32) * angp).astype(f32) + Rv).astype(f32)
        e3 = np.exp((((-b * 3.0) * dd) / ((f32(2) * Rv).astype(f64) + 1.0)) * t.astype(f64))
        e = np.where(stick1[None, :], e1, e2 if model == 2 else e3)
        out = (out.astype(f64) + f[k].astype(f64) * e).astype(f32)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        iso = np.where(ball1[None, :], np.exp(-b * dd), np.exp(da_ * np.log(dd / (dd + b * s2))))
    if f0 is not None:
        f0 = np.asarray(f0, f32)
        w = ((f32(1) - sumf).astype(f32) - f0).astype(f32)
        out = (out.astype(f64) + (f0.astype(f64) + w.astype(f64) * iso)).astype(f32)
    else:
        w = (f32(1) - sumf).astype(f32)
        out = (out.astype(f64) + w.astype(f64) * iso).astype(f32)
    return (out * S0[None, :]).astype(f32)
