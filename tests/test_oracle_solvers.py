"""Pins the numpy solver oracle (oracle/solvers.py) with textbook facts."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import solvers as SO  # noqa: E402


def dense(M):
    def apply(x, y):
        y += M @ x
    return apply


def spd_with_eigs(eigs, seed=0):
    n = len(eigs)
    Q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(n, n)))
    return (Q * eigs) @ Q.T


def test_cg_exact_termination_and_solution():
    # k distinct eigenvalues: CG terminates (to round-off) in k iterations
    eigs = np.repeat([1.0, 3.0, 7.0, 20.0], 10)
    A = spd_with_eigs(eigs)
    b = np.random.default_rng(1).normal(size=len(eigs))
    x, it, status, _ = SO.cg(dense(A), b, np.zeros_like(b), rtol=1e-10)
    assert status == 0
    assert it + 1 <= 5  # counter is the 0-based index of the converging iteration
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8)


def test_pcg_with_exact_preconditioner_converges_in_one_step():
    A = spd_with_eigs(np.linspace(1, 100, 30))
    Ainv = np.linalg.inv(A)
    b = np.random.default_rng(2).normal(size=30)
    x, it, status, _ = SO.cg(dense(A), b, np.zeros(30), precond=dense(Ainv), rtol=1e-10)
    assert status == 0 and it == 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10)


def test_power_method_matches_eigvalsh():
    A = spd_with_eigs(np.linspace(0.5, 10.0, 40), seed=3)
    lam = SO.power_method(dense(A), 40, 2000, 1e-12)
    assert abs(lam - 10.0) < 1e-6


def test_chebyshev3_is_the_scaled_chebyshev_polynomial():
    """For x0 = 0, b = A e: error after k steps is T_k((avg - l)/d) / T_k(avg/d) per
    eigencomponent, d = (lmax - lmin)/2 (three-term recurrence of the reference)."""
    eigs = np.linspace(0.01, 4.0, 25)
    A = np.diag(eigs)
    e = np.ones(25)
    for k in (1, 2, 3, 5):
        ch = SO.Chebyshev3(dense(A), 25, max_it=k, eig_max=4.0)
        x = ch(A @ e, np.zeros(25))
        lmax, lmin = 4.0, 0.06 * 4.0
        avg, d = (lmax + lmin) / 2, (lmax - lmin) / 2
        T = np.polynomial.chebyshev.Chebyshev.basis(k)
        expect = T((avg - eigs) / d) / T(avg / d)
        np.testing.assert_allclose(e - x, expect, rtol=1e-10, atol=1e-12)


def test_two_grid_error_propagator():
    """One V-cycle from x0 = 0 with linear Jacobi smoothing equals (I - E) A^-1 b with
    E = M (I - P Ac^-1 R A) M, M = I - w D^-1 A, Ac = R A P (textbook two-grid)."""
    n, nc = 31, 15
    A = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    P = np.zeros((n, nc))
    for j in range(nc):
        P[2 * j + 1, j] = 1
        P[2 * j, j] += 0.5
        P[2 * j + 2, j] += 0.5
    R = P.T
    Ac = R @ A @ P
    w = 2.0 / 3.0
    D = np.diag(A)
    smoother = SO.Jacobi(dense(A), D, w, steps=1)

    def coarse(b, x):
        x[:] = np.linalg.solve(Ac, b)
        return x

    mg = SO.Multigrid([(dense(A), smoother, None, lambda r: R @ r), (dense(Ac), coarse, lambda c: P @ c, None)],
                      max_it=1, atol=0)
    b = np.random.default_rng(4).normal(size=n)
    x = mg(b, np.zeros(n))
    M = np.eye(n) - w * (A / D[:, None])
    E = M @ (np.eye(n) - P @ np.linalg.solve(Ac, R @ A)) @ M
    np.testing.assert_allclose(x, (np.eye(n) - E) @ np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)


def test_jacobi_scaled_chebyshev_on_diagonal_matrix_matches_identity():
    """Chebyshev3 on D^-1 A with A = diag(d) is the unscaled smoother on A = I: the error of
    every component is multiplied by the same polynomial value."""
    rng = np.random.default_rng(5)
    n = 40
    dvec = rng.uniform(0.1, 100, n)
    A = lambda x, y: y.__iadd__(dvec * x)  # noqa: E731
    I = lambda x, y: y.__iadd__(x)  # noqa: E731
    b = rng.uniform(-1, 1, n)
    x0 = rng.uniform(-1, 1, n)
    xs = SO.Chebyshev3(A, n, max_it=3, scale_eig_max=1.02, eig_max_it=9, eig_tol=1e-12, inv_diag=1.0 / dvec)
    xi = SO.Chebyshev3(I, n, max_it=3, scale_eig_max=1.02, eig_max_it=9, eig_tol=1e-12)
    assert abs(xs.eig_max - 1.0) < 1e-12 and abs(xi.eig_max - 1.0) < 1e-12
    es = xs(b, x0.copy()) - b / dvec  # error w.r.t. the exact solution of A x = b
    ei = xi(b / dvec, x0.copy()) - b / dvec
    np.testing.assert_allclose(es, ei, rtol=1e-12, atol=1e-14)
    ratio = es / (x0 - b / dvec)
    assert np.ptp(ratio) < 1e-10 and abs(ratio[0]) < 1


def test_block_jacobi_on_block_diagonal_matrix_is_exact_with_omega_one():
    """For a block-diagonal SPD matrix, one block-Jacobi sweep with omega = 1 solves exactly; with
    a constrained dof the row is the identity (x_c += b_c - x_c)."""
    rng = np.random.default_rng(3)
    n = 7
    blocks = []
    for _ in range(n):
        M = rng.normal(size=(3, 3))
        blocks.append(M @ M.T + 3 * np.eye(3))
    Bd = np.zeros((3 * n, 3 * n))
    for i, M in enumerate(blocks):
        Bd[3 * i:3 * i + 3, 3 * i:3 * i + 3] = M
    A = lambda x, y: y.__iadd__(Bd @ x)  # noqa: E731
    sym = np.array([[M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]] for M in blocks])
    b = rng.uniform(-1, 1, 3 * n)
    x = SO.BlockJacobi(A, sym, omega=1.0, steps=1)(b, np.zeros(3 * n))
    np.testing.assert_allclose(Bd @ x, b, rtol=1e-12)
    # masked: dof 4 (block 1, row 1) becomes an identity row in the block used by the smoother
    mask = np.zeros(3 * n, np.uint8)
    mask[4] = 1
    bj = SO.BlockJacobi(A, sym, mask=mask, omega=1.0, steps=1)
    Bm = blocks[1].copy()
    Bm[1, :] = [0, 1, 0]
    np.testing.assert_allclose(bj.inv[1], np.linalg.inv(Bm), rtol=1e-12)
