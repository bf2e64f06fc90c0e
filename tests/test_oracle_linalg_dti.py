"""Pins of the oracle's linear algebra (the LU of CUDA/solver_mult_inverse.cu,
the pseudo-inverse, the symmetric eigen-decomposition) against numpy, and of the
tensor fit (DTI::linfit, diffmodels.cc:17-46) on noise-free tensor data."""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 2, 5, 8, 11, 12, 13, 14])
def test_lu_solve_spd_damped(orc, n):
    rng = np.random.default_rng(n)
    J = rng.standard_normal((3 * n, n)) * np.logspace(0, 4, n)
    A = 2 * J.T @ J + 0.1 * np.eye(n)
    b = rng.standard_normal(n)
    x = orc.lu_solve(A, b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-12)


def test_lu_solve_needs_pivoting(orc):
    rng = np.random.default_rng(5)
    for _ in range(50):
        n = int(rng.integers(2, 14))
        A = rng.standard_normal((n, n))
        A[0, 0] = 1e-14  # a zero-ish leading pivot forces a row interchange
        b = rng.standard_normal(n)
        x = orc.lu_solve(A, b)
        np.testing.assert_allclose(A @ x, b, rtol=1e-8, atol=1e-8 * np.abs(b).max())


def test_lu_solve_fp32_mode(orc):
    rng = np.random.default_rng(9)
    A = rng.standard_normal((6, 6)) + 6 * np.eye(6)
    b = rng.standard_normal(6)
    np.testing.assert_allclose(orc.lu_solve(A, b, prec=32), np.linalg.solve(A, b), rtol=1e-5, atol=1e-6)


def test_pinv_full_rank_and_rank_deficient(orc):
    rng = np.random.default_rng(2)
    A = rng.standard_normal((40, 7))
    np.testing.assert_allclose(orc.pinv(A), np.linalg.pinv(A), rtol=1e-9, atol=1e-12)
    B = rng.standard_normal((30, 3))
    B[:, 2] = B[:, 0] + B[:, 1]  # rank 2
    np.testing.assert_allclose(orc.pinv(B), np.linalg.pinv(B), rtol=1e-7, atol=1e-10)


def test_eig_against_eigh(orc):
    rng = np.random.default_rng(4)
    for n in (2, 3, 7):
        M = rng.standard_normal((n, n))
        S = M + M.T
        w, V = orc.eig(S)
        w2, V2 = np.linalg.eigh(S)
        np.testing.assert_allclose(w, w2, rtol=1e-10, atol=1e-12)
        for j in range(n):
            assert abs(abs(V[:, j] @ V2[:, j]) - 1) < 1e-8


def tensor_signal(S0, D, bvals, bvecs):
    return S0 * np.exp(-bvals * np.einsum("in,ij,jn->n", bvecs, D, bvecs))


def test_dti_noise_free_recovers_tensor(orc):
    rng = np.random.default_rng(11)
    n = 64
    g = rng.standard_normal((n, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    bvecs = np.ascontiguousarray(g.T)
    bvals = np.concatenate([np.zeros(4), np.full(n - 4, 1000.0)])
    for _ in range(20):
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        l = np.sort(rng.uniform(2e-4, 1.8e-3, 3))[::-1]
        D = Q @ np.diag(l) @ Q.T
        S0 = rng.uniform(500, 1500)
        y = tensor_signal(S0, D, bvals, bvecs)
        out = orc.dti(bvals, bvecs, y)[0]
        s0, md, fa, l1, l2, l3 = out[:6]
        assert s0 == pytest.approx(S0, rel=1e-9)
        np.testing.assert_allclose([l1, l2, l3], l, rtol=1e-7)
        assert md == pytest.approx(l.mean(), rel=1e-9)
        fa_ref = np.sqrt(1.5 * ((l - l.mean()) ** 2).sum() / (l ** 2).sum())
        assert fa == pytest.approx(fa_ref, rel=1e-7)
        v1 = out[6:9]
        assert abs(abs(v1 @ Q[:, 0]) - 1) < 1e-7
