"""CPU restatement of the CPU xfibres chain starts without nonlinear fitting
(test infrastructure only; only tests/ import this module).

* tensor_start: xfibresVoxelManager::initialise_tensor (xfibres.cc:546-572)
  from the tensor fit (DTI::linfit, restated in xfibres_oracle.cpp and pinned
  there): d = MD (model 2: d = d_std = 2 MD), d <= 0 -> 2e-3, all in float;
  S0 = the fit's S0; fibre 1 at the principal eigenvector with f = FA; the
  other fibres at Fibre's defaults th = pi/2, ph = 0, f = 0.01 (fibre.h:103-110);
  f0 at Multifibre's 0 (fibre.h:666).
* finish_stats: what Samples::record / finish_voxel (xfibres.cc:215-303) keep
  of a finished voxel in chain fibre order: float sums over the samples divided
  by their count, the dyadic tensor summed from float direction components,
  its principal eigenvector, and the number of fibres whose mean f exceeds 0.05.
  The eigenvector's sign is not fixed by the reference (NEWMAT's
  EigenValues); here, as on the device, its largest component is positive.
* neighbour_best: Samples::neighbour_initialise (xfibres.cc:306-357): scan x,
  then y, then z over the 3x3x3 block; a finished voxel whose supported-fibre
  count is at least maxfib and whose sum of mean f is larger than the best so
  far becomes the best and sets maxfib to its count; the start is taken from it
  only when maxfib ends above 1.  Outside the grid nothing is finished.
* expected_starts: the start every voxel gets in the CPU loop (xfibres.cc:898-921)
  given the final per-voxel state, since a voxel's state never changes after
  it finishes: when voxel v starts, exactly the voxels before it in mask order
  are finished.
* wavefronts: t = x + 2y + 4z, the order the device runs the voxels in.

The angles of cart2sph (FSL miscmaths, a dependency absent from the reference
tree) are atan2(y, x) and atan2(sqrt(x^2 + y^2), z), with th = ph = pi/2 for
the zero vector, the form the device uses.
"""
import numpy as np

f32 = np.float32


def cart2sph(v):
    v = np.asarray(v, np.float64)
    mag = np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    if mag == 0.0:
        return np.pi / 2, np.pi / 2
    return float(np.arctan2(np.sqrt(v[0] * v[0] + v[1] * v[1]), v[2])), float(np.arctan2(v[1], v[0]))


def n_params(model, K, f0):
    return 2 + 3 * K + (1 if model >= 2 else 0) + (1 if f0 else 0)


def tensor_start(model, K, f0, dti_row):
    """dti_row: the 15 values of oracle.dti for one voxel (s0, md, fa, l[3], v[3][3])
    -> chain start [S0, d, (d_std), (f, th, ph) x K, (f0)] as float32"""
    s0, md, fa = dti_row[0], dti_row[1], dti_row[2]
    v1 = np.asarray(dti_row[6:9])
    D = f32(md)
    if model >= 2:
        D = f32(D * f32(2))
    if D <= 0:
        D = f32(2e-3)
    add = 1 if model >= 2 else 0
    P = np.zeros(n_params(model, K, f0), f32)
    P[0] = f32(s0)
    P[1] = D
    if add:
        P[2] = D
    th, ph = cart2sph(v1)
    fib = [(f32(fa), f32(th), f32(ph))] + [(f32(0.01), f32(np.pi / 2), f32(0))] * (K - 1)
    for k, (f, t, p) in enumerate(fib):
        P[2 + add + 3 * k], P[3 + add + 3 * k], P[4 + add + 3 * k] = f, t, p
    return P


def neighbour_start(K, f0, best, mean_f, mean_d, mean_S0, dyad):
    """model-1 start from finished voxel `best`'s state"""
    P = np.zeros(n_params(1, K, f0), f32)
    P[0] = mean_S0[best]
    P[1] = mean_d[best]
    for k in range(K):
        th, ph = cart2sph(dyad[k, :, best])
        P[2 + 3 * k], P[3 + 3 * k], P[4 + 3 * k] = mean_f[k, best], f32(th), f32(ph)
    return P


def finish_stats(th, ph, f, d, S0):
    """raw chain-order samples of one voxel: th, ph, f [K][ns], d, S0 [ns]
    -> (been, mean_f [K], mean_d, mean_S0, dyad [K][3])"""
    th, ph, f = (np.asarray(a, f32) for a in (th, ph, f))
    K, ns = th.shape

    def fsum(x):
        acc = f32(0)
        for x_ in np.asarray(x, f32):
            acc = f32(acc + x_)
        return acc

    mean_d = f32(fsum(d) / f32(ns))
    mean_S0 = f32(fsum(S0) / f32(ns))
    mean_f = np.zeros(K, f32)
    dyad = np.zeros((K, 3))
    been = 0
    for k in range(K):
        mean_f[k] = f32(fsum(f[k]) / f32(ns))
        if float(mean_f[k]) > 0.05:
            been += 1
        st, ct, sp, cp = np.sin(th[k]), np.cos(th[k]), np.sin(ph[k]), np.cos(ph[k])
        x = np.stack([(st * cp).astype(f32), (st * sp).astype(f32), ct.astype(f32)]).astype(np.float64)
        T = x @ x.T
        w, V = np.linalg.eigh(T)
        e = V[:, int(np.argmax(w))]
        if e[int(np.argmax(np.abs(e)))] < 0:
            e = -e
        dyad[k] = e
    return been, mean_f, mean_d, mean_S0, dyad


def coords(gidx, dims):
    g = np.asarray(gidx, np.int64)
    nx, ny = dims[0], dims[1]
    return g % nx, (g // nx) % ny, g // (nx * ny)


def neighbour_best(v, gidx, dims, vol2mat, been_now, mean_f):
    """index of the voxel whose state starts voxel v, or -1 (tensor start);
    vol2mat: dict grid index -> mask index; been_now [nvox] as at v's start"""
    nx, ny, nz = dims
    x0, y0, z0 = (int(c[0]) for c in coords([gidx[v]], dims))
    K = mean_f.shape[0]
    maxfib, maxsf, best = 1, f32(0), -1
    for x in range(x0 - 1, x0 + 2):
        for y in range(y0 - 1, y0 + 2):
            for z in range(z0 - 1, z0 + 2):
                if not (0 <= x < nx and 0 <= y < ny and 0 <= z < nz):
                    continue
                vn = vol2mat.get(x + nx * (y + ny * z), -1)
                been = been_now[vn] if vn >= 0 else 0
                if been < maxfib:
                    continue
                sumf = f32(0)
                for k in range(K):
                    sumf = f32(sumf + mean_f[k, vn])
                if sumf > maxsf:
                    maxsf, maxfib, best = sumf, been, vn
    return best if maxfib > 1 else -1


def expected_starts(K, f0, gidx, dims, state, tensor_starts):
    """the start of every voxel in the CPU loop, from the final state of a run
    (been, mean_f, mean_d, mean_S0, dyad) and each voxel's tensor start;
    returns (starts [nvox][np], source [nvox]: the neighbour used or -1)"""
    nv = len(gidx)
    vol2mat = {int(g): i for i, g in enumerate(gidx)}
    starts = np.array(tensor_starts, f32, copy=True)
    src = np.full(nv, -1)
    been_now = np.zeros(nv, np.int64)
    for v in range(nv):
        b = neighbour_best(v, gidx, dims, vol2mat, been_now, state["mean_f"])
        if b >= 0:
            starts[v] = neighbour_start(K, f0, b, state["mean_f"], state["mean_d"], state["mean_S0"], state["dyad"])
            src[v] = b
        been_now[v] = state["been"][v]
    return starts, src


def wavefronts(gidx, dims):
    x, y, z = coords(gidx, dims)
    return x + 2 * y + 4 * z
