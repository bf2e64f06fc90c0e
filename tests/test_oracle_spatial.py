"""Pins of the CPU xfibres start restatement (oracle/spatial_oracle.py):
neighbour selection on hand-built neighbourhoods whose answers follow from
the text of Samples::neighbour_initialise (xfibres.cc:306-357), the mask-order
visibility of finished voxels, the wavefront order against brute force, the
finish statistics on closed-form sample sets, and the tensor start on
noise-free tensors (closed-form MD / FA / principal direction)."""
import itertools

import numpy as np
import pytest

from oracle import spatial_oracle as S

f32 = np.float32


def full_grid(dims):
    nx, ny, nz = dims
    gidx = np.arange(nx * ny * nz)
    return gidx, {int(g): i for i, g in enumerate(gidx)}


def lin(x, y, z, dims):
    return x + dims[0] * (y + dims[1] * z)


def centre_case(marks, K=2):
    """3x3x3 full grid, centre voxel 13; marks: {(dx,dy,dz): (been, [f per fibre])}"""
    dims = (3, 3, 3)
    gidx, v2m = full_grid(dims)
    been = np.zeros(27, np.int64)
    mean_f = np.zeros((K, 27), f32)
    for (dx, dy, dz), (b, fs) in marks.items():
        i = lin(1 + dx, 1 + dy, 1 + dz, dims)
        been[i] = b
        mean_f[:, i] = fs
    return S.neighbour_best(13, gidx, dims, v2m, been, mean_f), dims


def test_no_finished_or_single_fibre_neighbours_give_the_tensor_start():
    assert centre_case({})[0] == -1
    # finished neighbours with one supported fibre never start a voxel (maxfib must end above 1)
    assert centre_case({(-1, 0, 0): (1, [0.6, 0.01]), (0, 1, 1): (1, [0.9, 0.0])})[0] == -1


def test_larger_sum_of_f_wins_and_ties_keep_the_first():
    dims = (3, 3, 3)
    a, b = (-1, 0, 0), (1, 0, 0)  # a is scanned first (x outermost)
    best, _ = centre_case({a: (2, [0.3, 0.2]), b: (2, [0.4, 0.2])})
    assert best == lin(2, 1, 1, dims)
    best, _ = centre_case({a: (2, [0.4, 0.2]), b: (2, [0.3, 0.2])})
    assert best == lin(0, 1, 1, dims)
    best, _ = centre_case({a: (2, [0.3, 0.2]), b: (2, [0.3, 0.2])})  # equal sums: the strict > keeps a
    assert best == lin(0, 1, 1, dims)


def test_scan_order_x_outer_and_maxfib_filter():
    # P2 = (-1, +1, +1) comes before P1 = (+1, 0, -1) when x is the outer loop (the reference), after it when z is.
    # P2: 2 fibres, sum 0.9 -> best, maxfib 2; P1: 3 fibres but sum 0.4 < 0.9 -> not taken.
    dims = (3, 3, 3)
    best, _ = centre_case({(1, 0, -1): (3, [0.2, 0.2]), (-1, 1, 1): (2, [0.5, 0.4])})
    assert best == lin(0, 2, 2, dims)
    # a later candidate with fewer fibres than maxfib is skipped even with a larger sum
    best, _ = centre_case({(-1, -1, -1): (3, [0.2, 0.2]), (1, 1, 1): (2, [0.6, 0.3])})
    assert best == lin(0, 0, 0, dims)


def test_grid_edges_and_mask_holes():
    dims = (4, 3, 2)
    # mask without the voxel at (1, 0, 0)
    gidx = np.array([g for g in range(24) if g != lin(1, 0, 0, dims)])
    v2m = {int(g): i for i, g in enumerate(gidx)}
    been = np.zeros(len(gidx), np.int64)
    mean_f = np.zeros((2, len(gidx)), f32)
    j = v2m[lin(0, 1, 0, dims)]
    been[j], mean_f[:, j] = 2, [0.5, 0.3]
    v = v2m[lin(0, 0, 0, dims)]  # corner: only in-grid neighbours count
    assert S.neighbour_best(v, gidx, dims, v2m, been, mean_f) == j
    v = v2m[lin(3, 2, 1, dims)]  # far corner: (0,1,0) is not a neighbour
    assert S.neighbour_best(v, gidx, dims, v2m, been, mean_f) == -1


def test_expected_starts_see_only_voxels_earlier_in_mask_order():
    dims = (3, 1, 1)
    gidx = np.arange(3)
    K = 2
    state = dict(been=np.array([2, 2, 2]), mean_f=np.array([[0.3, 0.5, 0.7], [0.2, 0.2, 0.2]], f32),
                 mean_d=np.array([1e-3, 1.1e-3, 1.2e-3], f32), mean_S0=np.array([100, 110, 120], f32),
                 dyad=np.tile(np.array([0.0, 0.0, 1.0])[None, :, None], (K, 1, 3)))
    tstarts = np.full((3, S.n_params(1, K, False)), 7.0, f32)
    starts, src = S.expected_starts(K, False, gidx, dims, state, tstarts)
    # voxel 0 has nothing finished before it; voxel 1 sees only voxel 0 (voxel 2, with a larger sum, is later)
    assert list(src) == [-1, 0, 1]
    np.testing.assert_array_equal(starts[0], tstarts[0])
    assert starts[1][0] == f32(100) and starts[1][1] == f32(1e-3) and starts[1][2] == f32(0.3)
    assert starts[1][3] == f32(0.0)  # dyad along z: th = 0


@pytest.mark.parametrize("dims", [(1, 1, 5), (2, 3, 2), (3, 3, 3), (4, 2, 3), (5, 1, 2)])
def test_wavefronts_order_neighbours_as_the_mask_does(dims):
    nx, ny, nz = dims
    g = np.random.default_rng(sum(dims))
    gidx = np.sort(g.choice(nx * ny * nz, size=max(1, (nx * ny * nz * 3) // 4), replace=False))
    t = S.wavefronts(gidx, dims)
    x, y, z = S.coords(gidx, dims)
    for i, j in itertools.combinations(range(len(gidx)), 2):
        if max(abs(x[i] - x[j]), abs(y[i] - y[j]), abs(z[i] - z[j])) <= 1:
            # neighbours: earlier in mask order <=> earlier wavefront, never the same wavefront
            assert (gidx[i] < gidx[j]) == (t[i] < t[j]) and t[i] != t[j]


def test_finish_stats_closed_forms():
    ns = 40
    th = np.full((2, ns), f32(1.1))
    ph = np.full((2, ns), f32(-2.3))
    f = np.stack([np.full(ns, f32(0.05)), np.full(ns, f32(0.02))])
    d = np.full(ns, f32(1.3e-3))
    S0 = np.arange(ns, dtype=f32)
    been, mf, md, mS0, dyad = S.finish_stats(th, ph, f, d, S0)
    # the float sum of 40 x 0.05 rounds below 2, so the mean falls under 0.05: not supported
    assert mf[0] < f32(0.05) and been == 0
    assert abs(md - f32(1.3e-3)) <= np.spacing(f32(1.3e-3)) * 8
    assert mS0 == f32(sum(range(ns)) / ns)
    # one sample of float 0.05 lies above the double 0.05 the reference compares with: supported
    b1, mf1, _, _, _ = S.finish_stats(th[:, :1], ph[:, :1], f[:, :1], d[:1], S0[:1])
    assert mf1[0] == f32(0.05) and b1 == 1
    u = np.array([np.sin(1.1) * np.cos(-2.3), np.sin(1.1) * np.sin(-2.3), np.cos(1.1)])
    u = u if u[np.argmax(np.abs(u))] > 0 else -u
    np.testing.assert_allclose(dyad[0], u, atol=1e-6)
    # mixed directions: 30 samples along x, 10 along y -> principal direction x
    th2 = np.full((1, ns), f32(np.pi / 2))
    ph2 = np.concatenate([np.zeros(30, f32), np.full(10, f32(np.pi / 2))])[None]
    _, _, _, _, dy2 = S.finish_stats(th2, ph2, np.full((1, ns), f32(0.5)), d, S0)
    np.testing.assert_allclose(np.abs(dy2[0]), [1, 0, 0], atol=1e-6)


def test_tensor_start_on_noise_free_tensors(orc):
    lam = np.array([1.7e-3, 0.3e-3, 0.2e-3])
    ang = 0.4
    R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    D = R @ np.diag(lam) @ R.T
    g = np.random.default_rng(3)
    dirs = g.normal(size=(3, 60))
    dirs /= np.linalg.norm(dirs, axis=0)
    bvals = np.concatenate([np.zeros(4), np.full(60, 1000.0)])
    bvecs = np.concatenate([np.tile([[1.0], [0.0], [0.0]], (1, 4)), dirs], 1)
    S0 = 900.0
    sig = S0 * np.exp(-bvals * np.einsum("in,ij,jn->n", bvecs, D, bvecs))
    row = orc.dti(bvals, bvecs, sig[None])[0]
    md = lam.mean()
    fa = np.sqrt(1.5 * ((lam - md) ** 2).sum() / (lam ** 2).sum())
    for model, K, f0 in [(1, 2, False), (2, 3, True)]:
        P = S.tensor_start(model, K, f0, row)
        add = 1 if model == 2 else 0
        np.testing.assert_allclose(P[0], S0, rtol=1e-6)
        np.testing.assert_allclose(P[1], md * (2 if model == 2 else 1), rtol=1e-6)
        if model == 2:
            assert P[2] == P[1]
        np.testing.assert_allclose(P[2 + add], fa, rtol=1e-6)
        th, ph = P[3 + add], P[4 + add]
        v = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        assert abs(abs(v @ R[:, 0]) - 1) < 1e-6
        for k in range(1, K):
            assert (P[2 + add + 3 * k], P[3 + add + 3 * k], P[4 + add + 3 * k]) == (f32(0.01), f32(np.pi / 2), f32(0))
        if f0:
            assert P[-1] == 0
    # a signal that grows with b gives MD < 0: d falls back to 2e-3
    row2 = orc.dti(bvals, bvecs, (S0 * np.exp(bvals * 1e-4))[None])[0]
    assert row2[1] < 0
    assert S.tensor_start(1, 1, False, row2)[1] == f32(2e-3)
