"""Seeded synthetic diffusion data for the xfibres workloads.

Produces acquisitions (bvals, unit bvecs) and voxel signals drawn from the
ball-and-sticks family that xfibres fits, with Rician noise.  Everything is
seeded through numpy's Philox bit generator from the reference default seed
(8665904, xfibresoptions.h:128) plus a per-workload stream id, so a workload is
a pure function of its name and size.

Workloads (voxel counts follow the masks they are quoted on):
  "c0"  model-1 single shell: 5 b=0 + 64 dirs at b=1000, nf=2, ~2000 voxels,
        S0~U(800,1200), d~U(1e-3,2e-3), f1~U(0.3,0.6), second stick with
        p=0.5, f2~U(0.1,f1), crossing U(30,90) deg, Rician SNR 20.
  "c1"  multi-shell: 18 b=0 + 90 dirs at each of b=1000/2000/3000 (N=288),
        110x110x72 grid with an ellipsoid mask of semi-axes (48,48,26)
        (250,968 voxels); d~U(0.8e-3,2e-3), d_std=d*U(0.1,0.5), 1-3 fibres
        with weights summing to U(0.3,0.8), crossings U(30,90) deg, SNR 25.
  "c2_33", "c2_69", "c2_145", "c2_288": the mixed-acquisition volumes.
"""
from __future__ import annotations

import numpy as np

DEFAULT_SEED = 8665904

_STREAMS = {"c0": 0, "c1": 1, "c2_33": 21, "c2_69": 22, "c2_145": 23, "c2_288": 24, "c4": 4, "test": 99}


def rng(stream: str | int, sub: int = 0, seed: int = DEFAULT_SEED) -> np.random.Generator:
    sid = _STREAMS[stream] if isinstance(stream, str) else int(stream)
    return np.random.Generator(np.random.Philox(key=[seed, (sid << 16) | sub]))


def unit_vectors(g: np.random.Generator, n: int) -> np.ndarray:
    v = g.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def acquisition(n_b0: int, shells: list[tuple[float, int]], g: np.random.Generator):
    """bvals [N] and bvecs [3, N] (row-major x, y, z rows; b=0 volumes get (1,0,0))."""
    bvals = [np.zeros(n_b0)]
    vecs = [np.tile(np.array([[1.0, 0.0, 0.0]]), (n_b0, 1))]
    for b, nd in shells:
        bvals.append(np.full(nd, float(b)))
        vecs.append(unit_vectors(g, nd))
    return np.concatenate(bvals), np.ascontiguousarray(np.concatenate(vecs).T)


ACQ = {
    "c0": (5, [(1000.0, 64)]),
    "c1": (18, [(1000.0, 90), (2000.0, 90), (3000.0, 90)]),
    "c2_33": (3, [(1000.0, 30)]),
    "c2_69": (5, [(1000.0, 64)]),
    "c2_145": (10, [(1000.0, 45), (2000.0, 45), (3000.0, 45)]),
    "c2_288": (18, [(1000.0, 90), (2000.0, 90), (3000.0, 90)]),
    "c4": (18, [(1000.0, 90), (2000.0, 90), (3000.0, 90)]),
}


def ellipsoid_count(shape, axes) -> int:
    c = [(s - 1) / 2.0 for s in shape]
    x = (np.arange(shape[0]) - c[0]) / axes[0]
    y = (np.arange(shape[1]) - c[1]) / axes[1]
    z = (np.arange(shape[2]) - c[2]) / axes[2]
    r = x[:, None, None] ** 2 + y[None, :, None] ** 2 + z[None, None, :] ** 2
    return int((r <= 1.0).sum())


C1_NVOX = ellipsoid_count((110, 110, 72), (48, 48, 26))
C4_GRID, C4_AXES = (145, 174, 145), (60, 75, 48)
C4_NVOX = ellipsoid_count(C4_GRID, C4_AXES)
C4_R = 0.13  # zeppelin ratio of the model-3 truth (the R prior mean)
_C4_POS = None


def c4_positions() -> np.ndarray:
    """grid coordinates [C4_NVOX][3] of the high-resolution mask, in mask order"""
    global _C4_POS
    if _C4_POS is None:
        c = [(n - 1) / 2.0 for n in C4_GRID]
        x, y, z = np.meshgrid(*[np.arange(n) for n in C4_GRID], indexing="ij")
        r = ((x - c[0]) / C4_AXES[0]) ** 2 + ((y - c[1]) / C4_AXES[1]) ** 2 + ((z - c[2]) / C4_AXES[2]) ** 2
        _C4_POS = np.stack(np.nonzero(r <= 1.0), 1).astype(np.float64)
    return _C4_POS


def c4_snr(pos: np.ndarray) -> np.ndarray:
    """SNR field in [10, 30]: trilinear interpolation of seeded values on the 8 corners of the grid"""
    corner = rng("c4", 999).uniform(10.0, 30.0, (2, 2, 2))
    u = pos / (np.array(C4_GRID) - 1.0)
    out = np.zeros(len(pos))
    for i in (0, 1):
        for j in (0, 1):
            for k in (0, 1):
                w = (u[:, 0] if i else 1 - u[:, 0]) * (u[:, 1] if j else 1 - u[:, 1]) * (u[:, 2] if k else 1 - u[:, 2])
                out += w * corner[i, j, k]
    return out


def _perp_at_angle(u: np.ndarray, ang: np.ndarray, g: np.random.Generator) -> np.ndarray:
    """Unit vectors at angle `ang` (rad) from the unit vectors u (rows)."""
    r = g.standard_normal(u.shape)
    r -= (r * u).sum(1, keepdims=True) * u
    r /= np.linalg.norm(r, axis=1, keepdims=True)
    return np.cos(ang)[:, None] * u + np.sin(ang)[:, None] * r


def _sph(v: np.ndarray):
    th = np.arctan2(np.hypot(v[:, 0], v[:, 1]), v[:, 2])
    ph = np.arctan2(v[:, 1], v[:, 0])
    return th, ph


CHUNK = 16384


def voxels(name: str, nvox: int, sub: int = 0, with_truth: bool = True):
    """Returns dict(bvals, bvecs, data [nvox, N] float32 voxel-major, truth...).

    The acquisition depends on `name` only.  Voxels are drawn in fixed blocks
    of CHUNK, block c from Philox stream (name, 2000 + sub * 4096 + c): the
    result is a pure function of (name, nvox, sub), the first c*CHUNK voxels
    do not depend on nvox, and sub > 0 gives further voxels of the same
    distribution on the same acquisition (training sets)."""
    acq_g = rng(name, 1000)
    n_b0, shells = ACQ[name]
    bvals, bvecs = acquisition(n_b0, shells, acq_g)
    parts = []
    for c0 in range(0, nvox, CHUNK):
        extra = {}
        if name == "c4":  # SNR from the voxel's place in the high-resolution mask (wrapping for training sets)
            pos = c4_positions()
            idx = (np.arange(c0, min(nvox, c0 + CHUNK)) + sub * 104729) % len(pos)
            extra["snr"] = c4_snr(pos[idx])
        parts.append(_voxel_block(name, min(CHUNK, nvox - c0), rng(name, 2000 + sub * 4096 + c0 // CHUNK), bvals, bvecs,
                                  with_truth, **extra))
    out = {k: (np.concatenate([p[k] for p in parts]) if isinstance(parts[0][k], np.ndarray) else parts[0][k])
           for k in parts[0]}
    out.update(bvals=bvals, bvecs=bvecs)
    return out


def _voxel_block(name, nvox, g, bvals, bvecs, with_truth, snr=None):
    N = bvals.size
    K = 3
    S0 = g.uniform(800.0, 1200.0, nvox)
    dirs = np.zeros((nvox, K, 3))
    f = np.zeros((nvox, K))
    if name == "c0":
        d = g.uniform(1.0e-3, 2.0e-3, nvox)
        dstd = np.zeros(nvox)
        f[:, 0] = g.uniform(0.3, 0.6, nvox)
        two = g.uniform(size=nvox) < 0.5
        u = g.uniform(0.0, 1.0, nvox)
        f[:, 1] = np.where(two, 0.1 + u * (f[:, 0] - 0.1), 0.0)
        dirs[:, 0] = unit_vectors(g, nvox)
        ang = np.deg2rad(g.uniform(30.0, 90.0, nvox))
        dirs[:, 1] = _perp_at_angle(dirs[:, 0], ang, g)
        model, snr = 1, 20.0
        nfib_true = 1 + two.astype(np.int64)
    else:
        d = g.uniform(0.8e-3, 2.0e-3, nvox)
        dstd = d * g.uniform(0.1, 0.5, nvox)
        nfib_true = g.integers(1, 4, nvox)
        w = g.dirichlet(np.ones(3), nvox)
        w = np.where(np.arange(3)[None, :] < nfib_true[:, None], w, 0.0)
        w /= w.sum(1, keepdims=True)
        tot = g.uniform(0.3, 0.8, nvox)
        f = w * tot[:, None]
        dirs[:, 0] = unit_vectors(g, nvox)
        narrow = g.uniform(size=nvox) < 0.15 if name == "c4" else np.zeros(nvox, bool)
        for k in (1, 2):
            if name == "c4":  # 15% of voxels cross below 45 degrees
                ang = np.deg2rad(np.where(narrow, g.uniform(20.0, 45.0, nvox), g.uniform(45.0, 90.0, nvox)))
            else:
                ang = np.deg2rad(g.uniform(30.0, 90.0, nvox))
            base = dirs[:, k - 1] if k == 1 else dirs[:, 0]
            v = _perp_at_angle(base, ang, g)
            if k == 2:  # keep the third stick >= 30 deg from the second as well
                c = np.abs((v * dirs[:, 1]).sum(1))
                bad = c > np.cos(np.deg2rad(30.0))
                v[bad] = np.cross(dirs[bad, 0], dirs[bad, 1])
                v[bad] /= np.linalg.norm(v[bad], axis=1, keepdims=True)
            dirs[:, k] = v
        if name == "c4":
            dstd = np.minimum(dstd, 0.004)
            model = 3
        else:
            model, snr = 2, 25.0
    sig = clean_signal(model, bvals, bvecs, S0, d, dstd, f, dirs)
    sigma = S0 / snr
    n1 = g.standard_normal((nvox, N)) * sigma[:, None]
    n2 = g.standard_normal((nvox, N)) * sigma[:, None]
    data = np.sqrt((sig + n1) ** 2 + n2 ** 2).astype(np.float32)
    out = dict(data=data, model=model, snr=np.broadcast_to(np.asarray(snr, dtype=np.float64), (nvox,)).copy())
    if with_truth:
        th, ph = _sph(dirs.reshape(-1, 3))
        out.update(S0=S0, d=d, dstd=dstd, f=f, dirs=dirs, th=th.reshape(nvox, K), ph=ph.reshape(nvox, K),
                   nfib_true=nfib_true, clean=sig.astype(np.float32))
    return out


def clean_signal(model, bvals, bvecs, S0, d, dstd, f, dirs):
    """Noise-free ball-and-sticks signal [nvox, N] (float64).  bvals [N] and
    bvecs [3, N] are shared, or [nvox, N] and [nvox, 3, N] per voxel."""
    if np.ndim(bvecs) == 3:
        b = np.asarray(bvals)
        dp = np.einsum("vkc,vcn->vkn", dirs, bvecs)
    else:
        b = bvals[None, :]
        dp = np.einsum("vkc,cn->vkn", dirs, bvecs)
    a = dp * dp
    if model == 1:
        iso = np.exp(-d[:, None] * b)
        an = np.exp(-d[:, None, None] * b[:, None, :] * a)
    elif model == 3:
        # gamma-distributed isotropic diffusivity, zeppelin sticks with ratio C4_R
        # (the MCMC forms of CUDA/runmcmc_kernels.cu:67-95 with R fixed)
        s2 = dstd ** 2
        alpha = (d * d / s2)[:, None]
        iso = np.exp(np.log(d[:, None] / (d[:, None] + b * s2[:, None])) * alpha)
        R = C4_R
        an = np.exp(-3.0 * d[:, None, None] / (2 * R + 1) * b[:, None, :] * ((1 - R) * a + R))
    else:
        s2 = dstd ** 2
        alpha = (d * d / s2)[:, None]
        iso = np.exp(np.log(d[:, None] / (d[:, None] + b * s2[:, None])) * alpha)
        an = np.exp(np.log(d[:, None, None] / (d[:, None, None] + b[:, None, :] * a * s2[:, None, None])) * alpha[:, :, None])
    fs = f.sum(1)
    return S0[:, None] * ((1.0 - fs)[:, None] * iso + (f[:, :, None] * an).sum(1))


def grad_dev(g: np.random.Generator, nvox: int, sd: float = 0.03) -> np.ndarray:
    """per-voxel gradient-deviation values [nvox][9] (the grad_dev volume of a
    scanner with gradient nonlinearity): L(i, j) = grad[3j + i], entries ~ N(0, sd)"""
    return (g.standard_normal((nvox, 9)) * sd).astype(np.float32)


def corrected_acquisitions(bvals, bvecs, grad):
    """per-voxel (bvals [nvox][N], bvecs [nvox][3][N]) seen through the
    gradient deviations: v -> (I+L) v / |(I+L) v|, b -> b |(I+L) v|^2 for b > 0"""
    grad = np.asarray(grad, dtype=np.float64)
    L = grad.reshape(-1, 3, 3).transpose(0, 2, 1)           # L[v, i, j] = grad[v, 3j + i]
    M = np.eye(3)[None] + L
    w = np.einsum("vij,jn->vin", M, bvecs)
    mag = np.linalg.norm(w, axis=1)
    dw = bvals[None, :] > 0
    bc = np.where(dw[:, None, :], w / np.where(mag > 0, mag, 1.0)[:, None, :], bvecs[None])
    bv = np.where(dw, bvals[None, :] * mag ** 2, bvals[None, :])
    return bv, bc


def gradnonlin_voxels(name: str, nvox: int, sd: float = 0.03, clean_only: bool = False):
    """voxels of workload `name` whose signal was acquired through per-voxel
    gradient deviations (returned as "grad", [nvox][9] float32): same truth
    and noise draws as voxels(name, nvox), signal from the corrected
    acquisitions."""
    v = voxels(name, nvox)
    grad = grad_dev(rng(name, 3000), nvox, sd)
    model = v["model"]
    clean = np.empty((nvox, v["bvals"].size), np.float32)
    data = np.empty_like(clean)
    for c0 in range(0, nvox, CHUNK):
        sl = slice(c0, min(nvox, c0 + CHUNK))
        bv, bc = corrected_acquisitions(v["bvals"], v["bvecs"], grad[sl])
        sig = clean_signal(model, bv, bc, v["S0"][sl], v["d"][sl], v["dstd"][sl], v["f"][sl], v["dirs"][sl])
        g = rng(name, 3001 + c0 // CHUNK)
        sigma = (v["S0"][sl] / v["snr"][sl])[:, None]
        n1 = g.standard_normal(sig.shape) * sigma
        n2 = g.standard_normal(sig.shape) * sigma
        clean[sl] = sig
        data[sl] = clean[sl] if clean_only else np.sqrt((sig + n1) ** 2 + n2 ** 2)
    v["clean"], v["data"], v["grad"] = clean, data, grad
    return v
