"""Posterior agreement statistics between two finished chains on the same
voxels (the posterior parity measures): the supported-fibre count (mean f >
0.05, the ARD decision), the largest |mean f| difference over fibres, and the
largest angle between matched posterior-mean dyads on voxels with 2+
supported fibres on both sides.  Fibres are matched by the permutation of the
supported fibres that maximises the summed |cos| (the sort by summed f can
swap two fibres of similar weight between two chains)."""
from __future__ import annotations

import itertools

import numpy as np


def stats(a: dict, b: dict) -> dict:
    K = a["mean_f"].shape[0]
    na, nb = (a["mean_f"] > 0.05).sum(0), (b["mean_f"] > 0.05).sum(0)
    df = np.abs(a["mean_f"].astype(np.float64) - b["mean_f"]).max(0)
    two = np.nonzero((na >= 2) & (nb >= 2))[0]
    ang = np.zeros(len(two))
    for j, i in enumerate(two):
        A = [k for k in range(K) if a["mean_f"][k, i] > 0.05]
        B = [k for k in range(K) if b["mean_f"][k, i] > 0.05]
        best = None
        for perm in itertools.permutations(B, min(len(A), len(B))):
            cs = [abs(float(a["dyads"][p, :, i] @ b["dyads"][q, :, i])) for p, q in zip(A, perm)]
            if best is None or sum(cs) > sum(best):
                best = cs
        ang[j] = np.degrees(np.arccos(min(1.0, min(best))))
    return dict(n=len(df), n2=len(two), ard=float((na == nb).mean()), df=float((df <= 0.02).mean()),
                dyad=float((ang <= 2.0).mean()) if len(two) else 1.0,
                df_p95=float(np.percentile(df, 95)), dyad_p95=float(np.percentile(ang, 95)) if len(two) else 0.0)


BAR = dict(ard=0.98, df=0.95, dyad=0.95)


def tol(p: float, n: int) -> float:
    """three binomial standard errors of a proportion p measured on n voxels"""
    return 3.0 * np.sqrt(max(p * (1 - p), 0.01 * 0.99) / max(n, 1))


def fmt(s: dict) -> str:
    return (f"ARD {s['ard']:.3f}, |df|<=0.02 {s['df']:.3f} (p95 {s['df_p95']:.4f}), "
            f"dyads<=2deg {s['dyad']:.3f} (p95 {s['dyad_p95']:.2f}, n={s['n2']})")
