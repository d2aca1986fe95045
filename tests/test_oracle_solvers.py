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


def _textbook_pcg(A, b, precond, rtol, max_it):
    """Fletcher-Reeves PCG (beta = r_new.z_new / r_old.z_old): ||r_k|| / ||b|| per iteration."""
    x = np.zeros_like(b)
    r = b.copy()
    z = np.zeros_like(b)
    precond(r, z)
    p = z.copy()
    rtz = r @ z
    hist = []
    for k in range(max_it):
        Ap = A @ p
        alpha = rtz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        hist.append(np.linalg.norm(r) / np.linalg.norm(b))
        if hist[-1] < rtol:
            return x, k + 1, hist
        z = np.zeros_like(b)
        precond(r, z)
        rtz_new = r @ z
        p = z + (rtz_new / rtz) * p
        rtz = rtz_new
    return x, max_it, hist


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_flexible_pcg_equals_pcg_for_fixed_spd_preconditioner(seed):
    """With a fixed SPD preconditioner r_new.z_old = 0 in exact arithmetic: the flexible PCG's
    iterates are standard PCG's (residual histories to rounding, same count, same solution)."""
    n = 60
    A = spd_with_eigs(np.linspace(1, 100, n), seed=seed)
    b = np.random.default_rng(seed + 10).normal(size=n)
    d = np.diag(A) * np.random.default_rng(seed + 20).uniform(0.5, 2, n)

    def jac(r, z):
        z += r / d
    x, it, status, hist = SO.flexible_pcg(dense(A), b, np.zeros(n), jac, rtol=1e-10)
    xt, itt, ht = _textbook_pcg(A, b, jac, 1e-10, 1000)
    assert status == 0 and abs(it - itt) <= 1
    m = min(25, len(hist), len(ht))
    np.testing.assert_allclose(hist[:m], ht[:m], rtol=1e-8)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10  # the true residual is the stop rule


def test_flexible_pcg_converges_under_a_varying_preconditioner():
    """Jacobi with a fresh random scaling (1 +- 0.6) on every application is not a fixed map:
    the flexible beta still converges to the true tolerance, while the fixed-preconditioner beta
    (r_new.z_new / r_old.z_old, the same algorithm without the r_new.z_old term) stagnates."""
    n = 60
    A = spd_with_eigs(np.linspace(1, 100, n), seed=0)
    b = np.random.default_rng(10).normal(size=n)
    d = np.diag(A) * np.random.default_rng(20).uniform(0.5, 2, n)

    def varying(seed):
        rng = np.random.default_rng(seed)

        def P(r, z):
            z += r / d * (1 + 0.6 * rng.uniform(-1, 1, n))
        return P
    x, it, status, _ = SO.flexible_pcg(dense(A), b, np.zeros(n), varying(30), rtol=1e-10, max_it=1000)
    assert status == 0 and it < 400
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10
    _, itt, _ = _textbook_pcg(A, b, varying(30), 1e-10, 1000)
    assert itt == 1000  # without the flexible term: no convergence within 1000 iterations


def test_flexible_pcg_exact_preconditioner_and_residual_replacement():
    A = spd_with_eigs(np.linspace(1, 100, 30), seed=4)
    b = np.random.default_rng(5).normal(size=30)
    x, it, status, _ = SO.flexible_pcg(dense(A), b, np.zeros(30), dense(np.linalg.inv(A)), rtol=1e-10)
    assert status == 0 and it == 1
    d = np.diag(A)

    def jac(r, z):
        z += r / d
    x1, _, s1, _ = SO.flexible_pcg(dense(A), b, np.zeros(30), jac, rtol=1e-12)
    x2, _, s2, _ = SO.flexible_pcg(dense(A), b, np.zeros(30), jac, rtol=1e-12, residual_replacement=3)
    assert s1 == 0 and s2 == 0
    np.testing.assert_allclose(x2, x1, rtol=1e-9)
