"""numpy restatement of SFEM's solver shell (test infrastructure: the parity checker).

Operators are callables apply(x, y) that accumulate y += A x on float64 numpy vectors
(the oracle C library's apply functions, or dense matrices).  Each routine follows the
reference file:line it names; tests/test_oracle_solvers.py pins them against closed forms
(exact CG termination, eigvalsh, Chebyshev polynomials, the two-grid error propagator).
"""
import numpy as np


def _apply0(A, x):
    y = np.zeros_like(x)
    A(x, y)
    return y


def cg(A, b, x, rtol=1e-10, atol=1e-16, max_it=10000, precond=None):
    """ConjugateGradient::aux_apply_basic / aux_apply_precond (algebra/sfem_cg.hpp:301-445).
    Returns (x, iterations, status, history) with the reference's counter convention (the
    loop index at the converging iteration)."""
    x = x.copy()
    r = b - _apply0(A, x)
    hist = []
    if precond is None:
        rtr0 = r @ r
        if rtr0 == 0:
            return x, 0, 0, hist
        r_norm0 = np.sqrt(rtr0)
        p = r.copy()
        rtr = rtr0
        for it in range(max_it):
            Ap = _apply0(A, p)
            alpha = rtr / (p @ Ap)
            x += alpha * p
            r -= alpha * Ap
            rtr_new = r @ r
            beta = rtr_new / rtr
            rtr = rtr_new
            p = r + beta * p
            rn = np.sqrt(rtr_new)
            hist.append(rn / r_norm0)
            if rn < atol or rtr_new == 0 or rn / r_norm0 < rtol:
                return x, it, 0, hist
        return x, max_it, 1, hist
    if r @ r == 0:
        return x, 0, 0, hist
    z = _apply0(precond, r)
    p = z.copy()
    Ap = _apply0(A, p)
    rtz0 = r @ z
    rtz = rtz0
    if rtz == 0:
        return x, 0, 0, hist
    alpha = rtz / (p @ Ap)
    x += alpha * p
    r -= alpha * Ap
    for it in range(max_it):
        z = _apply0(precond, r)
        rtz_new = r @ z
        if rtz_new == 0:  # exact convergence; the next step would be 0/0 (guard of the build)
            return x, it, 0, hist
        beta = rtz_new / rtz
        rtz = rtz_new
        p = z + beta * p
        Ap = _apply0(A, p)
        alpha = rtz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        anorm = np.sqrt(abs(rtz))
        hist.append(anorm / np.sqrt(rtz0))
        if anorm < atol or anorm / np.sqrt(rtz0) < rtol:
            return x, it, 0, hist
    return x, max_it, 1, hist


def flexible_pcg(A, b, x, precond, rtol=1e-10, atol=1e-16, max_it=10000, residual_replacement=50):
    """The build's flexible PCG (the reference has none; its PCG is algebra/sfem_cg.hpp:365-445):
    beta = r_new.(z_new - z_old) / (r_old.z_old) with r_new.z_old taken before z is replaced; stops
    on the true relative residual ||b - A x|| / ||b|| (confirmed by recomputation), and replaces the
    recursive residual by b - A x every `residual_replacement` iterations.  Returns
    (x, iterations, status, history of relative residuals), iterations counted as completed steps.
    With a fixed SPD preconditioner r_new.z_old = 0 in exact arithmetic, so the iterates are those
    of standard PCG (tests/test_oracle_solvers.py pins it there and against dense solves)."""
    x = x.copy()
    bnorm = np.linalg.norm(b)
    ref = bnorm if bnorm > 0 else 1.0
    r = b - _apply0(A, x)
    rnorm = np.linalg.norm(r)
    hist = []
    if rnorm == 0 or rnorm / ref < rtol or rnorm < atol:
        return x, 0, 0, hist
    z = _apply0(precond, r)
    p = z.copy()
    rtz = r @ z
    for it in range(max_it):
        Ap = _apply0(A, p)
        alpha = rtz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        if residual_replacement > 0 and (it + 1) % residual_replacement == 0:
            r = b - _apply0(A, x)
        rtzold = r @ z
        rnorm = np.linalg.norm(r)
        hist.append(rnorm / ref)
        if rnorm / ref < rtol or rnorm < atol:
            r = b - _apply0(A, x)
            true_norm = np.linalg.norm(r)
            if true_norm / ref < rtol or true_norm < atol:
                hist[-1] = true_norm / ref
                return x, it + 1, 0, hist
            rtzold = r @ z
        z = _apply0(precond, r)
        rtz_new = r @ z
        beta = (rtz_new - rtzold) / rtz
        rtz = rtz_new
        p = z + beta * p
    return x, max_it, 1, hist


def power_method(A, n, max_it, tol, v0=None):
    """PowerMethod::max_eigen_value (algebra/sfem_PowerMethod.hpp:18-55): odd iteration
    count, stops when |lambda - lambda_old| <= tol on an even step."""
    max_it_odd = max_it + (max_it % 2 == 0)
    ev = np.ones(n) if v0 is None else v0.copy()
    work = _apply0(A, ev)
    work /= np.linalg.norm(work)
    vecs = [ev, work]
    lam = lam_old = 0.0
    for k in range(max_it_odd):
        w = vecs[k % 2]
        e = vecs[(k + 1) % 2]
        w[:] = _apply0(A, e)
        lam = np.linalg.norm(w)
        w /= lam
        if abs(lam - lam_old) <= tol and k % 2 == 0:
            break
        lam_old = lam
    return lam


class Chebyshev3:
    """Chebyshev3::precond_apply (algebra/sfem_Chebyshev3.hpp:183-242)."""

    def __init__(self, A, n, max_it=3, scale_eig_max=1.0, scale_eig_min=0.06, eig_max_it=1000, eig_tol=1e-6,
                 initial_guess_zero=False, eig_max=None, inv_diag=None, init=None):
        """inv_diag: run on D^-1 A (the build's Jacobi-scaled extension; the reference has none)."""
        self.A, self.n, self.max_it = A, n, max_it
        self.smax, self.smin = scale_eig_max, scale_eig_min
        self.zero = initial_guess_zero
        self.dinv = inv_diag
        def scaled(v, y):
            y += inv_diag * _apply0(A, v)

        Aop = A if inv_diag is None else scaled
        self.eig_max = power_method(Aop, n, eig_max_it, eig_tol, v0=init) if eig_max is None else eig_max

    def __call__(self, b, x):
        """x := smooth(b, x) in place."""
        lmax = self.eig_max * self.smax
        lmin = self.smin * lmax
        avg, diff = (lmin + lmax) / 2, (lmin - lmax) / 2
        alpha = 1.0 / avg
        d = 1.0 if self.dinv is None else self.dinv
        p = d * (-b.copy() if self.zero else _apply0(self.A, x) - b)
        x -= alpha * p
        for it in range(1, self.max_it):
            dea = diff * alpha
            beta = (0.5 if it == 1 else 0.25) * dea * dea
            alpha = 1.0 / (avg - beta / alpha)
            p = beta * p + d * (_apply0(self.A, x) - b)
            x -= alpha * p
        return x


class Jacobi:
    """StationaryIteration with scaled Jacobi (algebra/sfem_Stationary.hpp:50-80)."""

    def __init__(self, A, diag, omega=1.0, steps=3):
        self.A, self.inv, self.steps = A, omega / diag, steps

    def __call__(self, b, x):
        for _ in range(self.steps):
            x += self.inv * (b - _apply0(self.A, x))
        return x


class BlockJacobi:
    """StationaryIteration with ShiftableBlockSymJacobi (algebra/sfem_ShiftableJacobi.hpp set_diag,
    algebra/openmp/sfem_openmp_ShiftableJacobi.hpp): symmetric blocks [n][6] -> full 3x3, rows of
    constrained dofs replaced by unit rows, inverted, scaled by omega; x += Binv (b - A x)."""

    def __init__(self, A, sym_blocks, mask=None, omega=1.0 / 3, steps=3):
        d = np.asarray(sym_blocks, np.float64).reshape(-1, 6)
        B = np.stack([d[:, [0, 1, 2]], d[:, [1, 3, 4]], d[:, [2, 4, 5]]], axis=1)  # [n][3][3]
        if mask is not None:
            m = np.asarray(mask).reshape(-1, 3).astype(bool)
            for r in range(3):
                B[m[:, r], r, :] = 0
                B[m[:, r], r, r] = 1
        self.inv = omega * np.linalg.inv(B)
        self.A, self.steps = A, steps

    def __call__(self, b, x):
        for _ in range(self.steps):
            r = (b - _apply0(self.A, x)).reshape(-1, 3)
            x += np.einsum("bij,bj->bi", self.inv, r).ravel()
        return x


class Multigrid:
    """Multigrid::cycle (algebra/sfem_Multigrid.hpp:242-346).  levels: list of
    (A, smoother(b, x) -> x, prolongation(c) -> f or None, restriction(f) -> c or None);
    the coarsest smoother is the coarse solver (called with x = 0)."""

    def __init__(self, levels, max_it=10, atol=1e-10, cycle_type=1):
        self.levels, self.max_it, self.atol, self.cycle_type = levels, max_it, atol, cycle_type
        self.history = []

    def _cycle(self, l, b, x):
        A, smoother, _, R = self.levels[l]
        if l == len(self.levels) - 1:
            x[:] = 0
            smoother(b, x)
            return False
        P = self.levels[l + 1][2]
        for _ in range(self.cycle_type):
            smoother(b, x)
            r = b - _apply0(A, x)
            if l == 0:
                nr = np.linalg.norm(r)
                self.history.append(nr)
                if nr < self.atol:
                    return True
            rc = R(r)
            xc = np.zeros_like(rc)
            self._cycle(l + 1, rc, xc)
            x += P(xc)
            smoother(b, x)
        return False

    def __call__(self, b, x):
        self.history = []
        for _ in range(self.max_it):
            if self._cycle(0, b, x):
                break
        return x
