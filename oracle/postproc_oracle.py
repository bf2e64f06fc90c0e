"""CPU restatement of the bedpostx post-processing (test infrastructure only).

What the reference's bedpostx_postproc_gpu.sh does after merge_parts_gpu,
restated with numpy on [..., ns] sample arrays of the masked voxels:
  tmean            fslmaths <merged> -Tmean            (bedpostx_postproc_gpu.sh:23-25)
  dyadic_vectors   make_dyadic_vectors th ph mask out  (make_dyadic_vectors.cc:59-96;
                   mean of v v^T over samples, principal eigenvector, dispersion
                   1 - largest |eigenvalue|)
  maskdyads        fslmaths f -thr 0.05 -bin -mul dyads (maskdyads:18-24)
  modf             dyads_thr x (mean_f_k / mean_f_1)    (bedpostx_postproc_gpu.sh:30-32;
                   fslmaths -div gives 0 where the divisor is 0)
  fsum             sum of mean_f_k                      (bedpostx_postproc_gpu.sh:40-47)
Only tests/ import this module; the product never does.
"""
import numpy as np


def tmean(samples):
    """mean over the last axis, accumulated in double, rounded to float32"""
    return (np.asarray(samples, dtype=np.float64).mean(-1)).astype(np.float32)


def dyadic_vectors(th, ph):
    """th, ph [nvox][ns] -> (dyads [nvox][3] (sign arbitrary), dispersion [nvox])"""
    th = np.asarray(th, dtype=np.float32)
    ph = np.asarray(ph, dtype=np.float32)
    d = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1).astype(np.float64)
    T = np.einsum("vsi,vsj->vij", d, d) / th.shape[-1]
    w, V = np.linalg.eigh(T)
    return V[:, :, -1], 1.0 - np.abs(w).max(-1)


def maskdyads(dyads, mean_f, thr=0.05):
    """dyads [nvox][3] kept where mean_f >= thr, zero elsewhere (float32)"""
    keep = (np.asarray(mean_f, dtype=np.float32) >= np.float32(thr)).astype(np.float32)
    return (np.asarray(dyads, dtype=np.float32) * keep[:, None]).astype(np.float32)


def modf(dyads_thr, mean_fk, mean_f1):
    mean_fk = np.asarray(mean_fk, dtype=np.float32)
    mean_f1 = np.asarray(mean_f1, dtype=np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(mean_f1 != 0, mean_fk / np.where(mean_f1 != 0, mean_f1, np.float32(1)), np.float32(0))
    return (np.asarray(dyads_thr, dtype=np.float32) * r.astype(np.float32)[:, None]).astype(np.float32)


def fsum(mean_fs):
    s = np.zeros_like(np.asarray(mean_fs[0], dtype=np.float32))
    for m in mean_fs:
        s = (s + np.asarray(m, dtype=np.float32)).astype(np.float32)
    return s
