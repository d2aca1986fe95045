"""GPU solver parity (CG, Chebyshev3/PowerMethod, one V-cycle, MG-PCG, mixed-precision
flexible CG) against the numpy solver oracle (oracle/solvers.py) driving the C oracle."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

import sfem_amd as S  # noqa: E402
from sfem_amd import gmg  # noqa: E402

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import solvers as SO  # noqa: E402


def dev(a, dtype=np.float64):
    return S.DeviceArray.from_numpy(np.asarray(a), dtype)


def host(t):
    return t.numpy()


def zeros(n, dtype=np.float64):
    return S.DeviceArray(n, dtype)


def true_residual(A, b, x):
    r = zeros(b.n, b.dtype)
    A.apply(x, r)
    bb = b.numpy()
    return float(np.linalg.norm(bb - r.numpy()) / np.linalg.norm(bb))


def oracle_constrained(apply, nodes, bs=1):
    dofs = (nodes[:, None].astype(np.int64) * bs + np.arange(bs)).ravel()

    def f(x, y):
        apply(x, y)
        y[dofs] = x[dofs]
    return f


def test_cg_tet4_poisson_matches_oracle_iterations():
    """Config 1 shape: 32^3 cubes x 6 tets, Dirichlet x = 0, f = 1, CG to 1e-10."""
    el, pts = O.box_mesh(33, 33, 33, tet4=True)
    fff = O.tet4_fff(el, pts)
    N = pts.shape[1]
    nodes = np.where(pts[0] == 0)[0].astype(np.int32)
    b = np.ones(N)
    b[nodes] = 0
    A_o = oracle_constrained(lambda x, y: O.lib.oracle_tet4_laplacian_apply(el.shape[1], el, fff, x, y), nodes)
    xo, it_o, st_o, _ = SO.cg(A_o, b, np.zeros(N), rtol=1e-10, atol=1e-16, max_it=10000)
    assert st_o == 0
    A = S.constrained(S.laplacian(S.Mesh(S.TET4, 1, el, N), fff), nodes)
    solver = S.cg(A, rtol=1e-10, atol=1e-16, max_it=10000)
    x = zeros(N)
    solver.apply(dev(b), x)
    print(f"tet4 CG: gpu {solver.iterations} its, oracle {it_o} its")
    assert solver.last_status == 0
    assert abs(solver.iterations - it_o) <= 1
    assert np.linalg.norm(host(x) - xo) / np.linalg.norm(xo) < 1e-8
    assert true_residual(A, dev(b), x) < 2e-10


def _ss_poisson(L, n, coeff_seed=None):
    mel, mpts = O.box_mesh(n + 1, n + 1, n + 1)
    el, pts = O.sshex8_refine(L, mel, mpts)
    fff = O.hex8_fff(mel, mpts)
    N = pts.shape[1]
    nodes = np.where((pts.min(0) < 1e-6) | (pts.max(0) > 1 - 1e-6))[0].astype(np.int32)
    coeff = None
    if coeff_seed is not None:
        coeff = (10 ** np.random.default_rng(coeff_seed).uniform(0, 3, mel.shape[1])).astype(np.float32)
    return mel, mpts, el, pts, fff, N, nodes, coeff


def test_power_method_and_chebyshev_parity():
    L = 4
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 3, coeff_seed=1)
    E = mel.shape[1]
    A_o = oracle_constrained(lambda x, y: O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, coeff, x, y), nodes)
    v0 = np.ones(N)
    v0[nodes] = 0
    lam_o = SO.power_method(A_o, N, 10, 1e-5, v0=v0)
    A = S.constrained(S.laplacian(S.Mesh(S.SSHEX8, L, el, N), fff, coeff), nodes)
    ch = S.cheb3(A, max_it=3, scale_eig_max=1.02, eigen_solver_max_it=10, eigen_solver_tol=1e-5, init_vector=dev(v0))
    lam = S.cheb3_eig_max(ch)
    assert abs(lam - lam_o) < 1e-10 * lam_o
    rng = np.random.default_rng(0)
    b = rng.uniform(-1, 1, N)
    b[nodes] = 0
    x0 = rng.uniform(-1, 1, N)
    x0[nodes] = 0
    xo = SO.Chebyshev3(A_o, N, max_it=3, scale_eig_max=1.02, eig_max=lam_o)(b, x0.copy())
    x = dev(x0)
    ch.apply(dev(b), x)
    S.synchronize()
    assert np.linalg.norm(host(x) - xo) / np.linalg.norm(xo) < 1e-12


def test_jacobi_scaled_chebyshev_parity():
    """Chebyshev3 on D^-1 A (extension) against the oracle with the same scaling."""
    L = 4
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 3, coeff_seed=2)
    E = mel.shape[1]
    A_o = oracle_constrained(lambda x, y: O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, coeff, x, y), nodes)
    d = np.zeros(N)
    O.lib.oracle_sshex8_laplacian_diag(L, E, el, fff, coeff, d)
    d[nodes] = 1
    v0 = np.ones(N)
    v0[nodes] = 0
    A = S.constrained(S.laplacian(S.Mesh(S.SSHEX8, L, el, N), fff, coeff), nodes)
    dd = zeros(N)
    A.diag(dd)
    assert np.linalg.norm(host(dd) - d) < 1e-12 * np.linalg.norm(d)
    ch_o = SO.Chebyshev3(A_o, N, max_it=3, scale_eig_max=1.02, eig_max_it=10, eig_tol=1e-5, inv_diag=1.0 / d,
                         init=v0)
    ch = S.cheb3(A, max_it=3, scale_eig_max=1.02, eigen_solver_max_it=10, eigen_solver_tol=1e-5, init_vector=dev(v0),
                 inv_diag=dev(1.0 / host(dd)))
    assert abs(S.cheb3_eig_max(ch) - ch_o.eig_max) < 1e-10 * ch_o.eig_max
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, N)
    b[nodes] = 0
    x0 = rng.uniform(-1, 1, N)
    x0[nodes] = 0
    xo = ch_o(b, x0.copy())
    x = dev(x0)
    ch.apply(dev(b), x)
    S.synchronize()
    assert np.linalg.norm(host(x) - xo) / np.linalg.norm(xo) < 1e-12


@pytest.mark.parametrize("scaled", [False, True])
def test_lanczos_eig_max_matches_dense_spectrum(scaled):
    """Chebyshev3 with its bound from Lanczos (extension): lambda_max of the constrained sshex8
    Laplacian (L=4 on 3^3, coefficients over three decades), or of D^-1 A, against the dense
    spectrum of the free block (numpy eigvalsh of the oracle's matrix, symmetrically scaled);
    the power method's estimate lies below it; the smoother with that bound matches the oracle's
    Chebyshev3 step for step."""
    L = 4
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 3, coeff_seed=4)
    E = mel.shape[1]
    A_o = oracle_constrained(lambda x, y: O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, coeff, x, y), nodes)
    free = np.setdiff1d(np.arange(N), nodes)
    M = np.zeros((free.size, free.size))
    for j, c in enumerate(free):
        e = np.zeros(N)
        e[c] = 1
        y = np.zeros(N)
        A_o(e, y)
        M[:, j] = y[free]
    d = np.zeros(N)
    O.lib.oracle_sshex8_laplacian_diag(L, E, el, fff, coeff, d)
    d[nodes] = 1
    if scaled:
        s = 1 / np.sqrt(d[free])
        M = s[:, None] * M * s[None, :]
    lam_dense = float(np.linalg.eigvalsh(0.5 * (M + M.T))[-1])
    v0 = np.ones(N)
    v0[nodes] = 0
    A = S.constrained(S.laplacian(S.Mesh(S.SSHEX8, L, el, N), fff, coeff), nodes)
    inv = dev(1.0 / d) if scaled else None
    ch = S.cheb3(A, max_it=3, scale_eig_max=1.02, eigen_solver_tol=1e-13, init_vector=dev(v0), inv_diag=inv,
                 lanczos_steps=free.size)
    lam = S.cheb3_eig_max(ch)
    steps = S.cheb3_eigen_iterations(ch)
    pw = S.cheb3(A, max_it=3, scale_eig_max=1.02, eigen_solver_max_it=1000, eigen_solver_tol=1e-5,
                 init_vector=dev(v0), inv_diag=inv)
    lam_pw = S.cheb3_eig_max(pw)
    print(f"lambda_max dense {lam_dense:.15g}, Lanczos {lam:.15g} ({steps} steps), power {lam_pw:.15g} "
          f"({S.cheb3_eigen_iterations(pw)} its)")
    assert abs(lam - lam_dense) < 1e-10 * lam_dense
    assert steps < 200
    assert lam_pw <= lam_dense * (1 + 1e-12)
    rng = np.random.default_rng(5)
    b = rng.uniform(-1, 1, N)
    b[nodes] = 0
    x0 = rng.uniform(-1, 1, N)
    x0[nodes] = 0
    xo = SO.Chebyshev3(A_o, N, max_it=3, scale_eig_max=1.02, eig_max=lam,
                       inv_diag=(1.0 / d) if scaled else None)(b, x0.copy())
    x = dev(x0)
    ch.apply(dev(b), x)
    S.synchronize()
    assert np.linalg.norm(host(x) - xo) / np.linalg.norm(xo) < 1e-12


def _oracle_hierarchy(L, mel, mpts, el, pts, fff, coeff, nodes, lams):
    """Oracle V-cycle mirroring sfem_amd.gmg.SSGMG (same eig_max values)."""
    E = mel.shape[1]
    levels = gmg.derefinement_levels(L)
    conns = [el] + [O.derefine(L, l, el) for l in levels[1:]]
    out = []
    for i, (l, c) in enumerate(zip(levels, conns)):
        n = c.max() + 1
        bc = nodes[nodes < n]
        if l == 1:
            h8 = O.to_hex8(1, c)
            ap = (lambda h8: lambda x, y: O.lib.oracle_hex8_laplacian_apply(E, h8, fff, coeff, x, y))(h8)
        else:
            ap = (lambda l, c: lambda x, y: O.lib.oracle_sshex8_laplacian_apply(l, E, c, fff, coeff, x, y))(l, c)
        A = oracle_constrained(ap, bc)
        if i < len(levels) - 1:
            sm = SO.Chebyshev3(A, n, max_it=3, scale_eig_max=1.02, eig_max=lams[i])
        else:
            d = np.zeros(n)
            O.lib.oracle_hex8_laplacian_diag(E, O.to_hex8(1, c), fff, coeff, d)
            d[bc] = 1

            def sm(b, x, A=A, d=d):
                x[:] = SO.cg(A, b, np.zeros_like(b), rtol=1e-12, atol=1e-30,
                             precond=lambda r, z: z.__iadd__(r / d))[0]
                return x
        out.append([A, sm, None, None, n, bc, c, l])
    for i in range(len(out) - 1):
        f, co = out[i], out[i + 1]
        cnt = O.incidence(f[6], f[4])

        def R(r, f=f, co=co, cnt=cnt):
            rc = np.zeros(co[4])
            if co[7] == 1:
                O.lib.oracle_sshex8_hierarchical_restriction(f[7], E, f[6], cnt, 1, r, rc)
            else:
                O.lib.oracle_sshex8_restrict(E, f[7], f[6], cnt, co[7], co[6], 1, r, rc)
            rc[co[5]] = 0
            return rc

        def P(xc, f=f, co=co):
            xf = np.zeros(f[4])
            if co[7] == 1:
                O.lib.oracle_sshex8_hierarchical_prolongation(f[7], E, f[6], 1, xc, xf)
            else:
                O.lib.oracle_sshex8_prolongate(E, co[7], co[6], f[7], f[6], 1, xc, xf)
            xf[f[5]] = 0
            return xf
        f[3] = R
        co[2] = P
    return SO.Multigrid([tuple(o[:4]) for o in out], max_it=1, atol=0)


def test_single_vcycle_matches_oracle():
    L = 4
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 3, coeff_seed=2)
    h = gmg.SSGMG(S.Mesh(S.SSHEX8, L, el, N), "laplacian", fff, nodes, coeff=coeff, coarse_rtol=1e-12)
    lams = [S.cheb3_eig_max(lv.smoother) for lv in h.levels[:-1]]
    mg_o = _oracle_hierarchy(L, mel, mpts, el, pts, fff, coeff, nodes, lams)
    rng = np.random.default_rng(3)
    b = rng.uniform(0, 1, N)
    b[nodes] = 0
    xo = mg_o(b, np.zeros(N))
    x = zeros(N)
    h.mg.apply(dev(b), x)
    S.synchronize()
    err = np.linalg.norm(host(x) - xo) / np.linalg.norm(xo)
    print(f"one V-cycle: relative difference to the oracle {err:.2e}")
    assert err < 1e-9


@pytest.mark.parametrize("cycle_type", [1, 2])
def test_multigrid_as_solver_matches_oracle(cycle_type):
    """Multigrid as the solver (5 cycles, V or W): the iterate and the finest-level residual norm of
    every cycle (computed after the pre-smoothing, algebra/sfem_Multigrid.hpp:242-346) against the
    oracle's cycle."""
    L = 4
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 3, coeff_seed=4)
    h = gmg.SSGMG(S.Mesh(S.SSHEX8, L, el, N), "laplacian", fff, nodes, coeff=coeff, coarse_rtol=1e-12,
                  cycle_type=cycle_type, max_it=5, atol=0.0)
    lams = [S.cheb3_eig_max(lv.smoother) for lv in h.levels[:-1]]
    mg_o = _oracle_hierarchy(L, mel, mpts, el, pts, fff, coeff, nodes, lams)
    mg_o.max_it, mg_o.cycle_type = 5, cycle_type
    b = np.random.default_rng(6).uniform(-1, 1, N)
    b[nodes] = 0
    xo = mg_o(b, np.zeros(N))
    x = zeros(N)
    h.mg.apply(dev(b), x)
    S.synchronize()
    hist = h.mg.history()
    err = np.linalg.norm(host(x) - xo) / np.linalg.norm(xo)
    herr = np.max(np.abs(hist - np.array(mg_o.history)) / np.array(mg_o.history))
    print(f"{'V' if cycle_type == 1 else 'W'}-cycle solver: iterate {err:.2e}, residual history {herr:.2e}, "
          f"residual reduction {hist[-1] / hist[0]:.2e} over {len(hist) - 1} norms (jumping coefficients, unscaled "
          f"Chebyshev3)")
    # the cycle repeats cycle_type times on every level, the finest included (one norm each)
    assert len(hist) == 5 * cycle_type and len(mg_o.history) == 5 * cycle_type
    assert err < 1e-9 and herr < 1e-8


@pytest.mark.parametrize("coeff_seed", [None, 5])
def test_mg_pcg_poisson_converges_and_matches_oracle(coeff_seed):
    L = 8
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 4, coeff_seed)
    h = gmg.SSGMG(S.Mesh(S.SSHEX8, L, el, N), "laplacian", fff, nodes, coeff=coeff, coarse_rtol=1e-12)
    lams = [S.cheb3_eig_max(lv.smoother) for lv in h.levels[:-1]]
    b = np.random.default_rng(0).uniform(0, 1, N)
    b[nodes] = 0
    solver = S.cg(h.fine.A, h.mg, rtol=1e-10, max_it=200)
    x = zeros(N)
    solver.apply(dev(b), x)
    mg_o = _oracle_hierarchy(L, mel, mpts, el, pts, fff, coeff, nodes, lams)
    xo, it_o, st_o, _ = SO.cg(h_apply := mg_o.levels[0][0], b, np.zeros(N), rtol=1e-10, max_it=200,
                              precond=lambda r, z: z.__iadd__(mg_o(r, np.zeros_like(r))))
    print(f"MG-PCG: gpu {solver.iterations} its, oracle {it_o} its")
    assert solver.last_status == 0 and st_o == 0
    assert abs(solver.iterations - it_o) <= 1
    assert true_residual(h.fine.A, dev(b), x) < 1e-8
    assert np.linalg.norm(host(x) - xo) / np.linalg.norm(xo) < 1e-7


def test_mixed_precision_flexible_cg_reaches_fp64_tolerance():
    L = 8
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 4, coeff_seed=7)
    mesh = S.Mesh(S.SSHEX8, L, el, N)
    h64 = gmg.SSGMG(mesh, "laplacian", fff, nodes, coeff=coeff)
    h32 = gmg.SSGMG(mesh, "laplacian", fff, nodes, coeff=coeff, real_type=S.FLOAT32)
    P = S.precision_bridge(h32.mg)
    solver = S.cg(h64.fine.A, P, flexible=True, rtol=1e-10, max_it=200)
    b = dev(np.where(np.isin(np.arange(N), nodes), 0.0, np.random.default_rng(0).uniform(0, 1, N)))
    x = zeros(N)
    solver.apply(b, x)
    res = true_residual(h64.fine.A, b, x)
    print(f"mixed MG-FCG: {solver.iterations} its, true FP64 residual {res:.2e}")
    assert solver.last_status == 0 and res < 1e-10


@pytest.mark.parametrize("precond", ["vcycle", "vcycle_graph", "jacobi"])
def test_mixed_flexible_cg_keeps_z_in_fp32_same_iterates(precond):
    """A flexible CG around an FP32 preconditioner behind a precision bridge keeps z in FP32 (option
    mixed_cg, the default): the CG update writes the FP32 residual and the dots and the direction
    update read z in FP32.  Its iterates are those of the generic flexible CG on the bridge's
    (double) z, bit for bit: iteration count, residual history and solution.  The FP32 Jacobi case
    runs past the residual replacement (every 50 iterations)."""
    L = 8
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 4, coeff_seed=7)
    mesh = S.Mesh(S.SSHEX8, L, el, N)
    h64 = gmg.SSGMG(mesh, "laplacian", fff, nodes, coeff=coeff)
    # (a graph replay needs a V-cycle without host synchronisation: the dense coarse solve)
    h32 = gmg.SSGMG(mesh, "laplacian", fff, nodes, coeff=coeff, real_type=S.FLOAT32, coarse="dense")
    if precond == "jacobi":
        d32 = zeros(N, np.float32)
        h32.fine.A.diag(d32)
        S.synchronize()
        P = S.precision_bridge(S.inverse_diagonal(d32))
    else:
        P = S.precision_bridge(h32.mg)
        if precond == "vcycle_graph":
            P = S.graph_op(P)
    b = dev(np.where(np.isin(np.arange(N), nodes), 0.0, np.random.default_rng(0).uniform(0, 1, N)))
    runs = {}
    for mode in (1, 0):
        S.set_option("mixed_cg", mode)
        try:
            solver = S.cg(h64.fine.A, P, flexible=True, rtol=1e-10, max_it=1000)
        finally:
            S.set_option("mixed_cg", 1)
        sols = []
        for _ in range(2):  # the second solve replays the graphs
            x = zeros(N)
            solver.apply(b, x)
            S.synchronize()
            sols.append(host(x))
        runs[mode] = (solver.iterations, solver.last_status, solver.history(), sols)
    its, st, hist, sols = runs[1]
    print(f"{precond}: {its} iterations (generic flexible CG {runs[0][0]})")
    assert st == 0 and runs[0][1] == 0
    assert its == runs[0][0]
    if precond == "jacobi":
        assert its > 50
    assert np.array_equal(hist, runs[0][2])
    for a, g in zip(sols, runs[0][3]):
        assert np.array_equal(a, g)
    assert true_residual(h64.fine.A, b, dev(sols[1])) < 1e-10


def test_elasticity_mg_pcg_converges():
    L = 4
    mel, mpts = O.box_mesh(4, 4, 4)
    el, pts = O.sshex8_refine(L, mel, mpts)
    adj, det = O.hex8_adj_det(mel, mpts)
    N = pts.shape[1]
    nodes = np.where(pts[2] < 1e-6)[0].astype(np.int32)  # clamped at z = 0
    mu, lam = O.lame(1.0, 0.3)
    coeff = np.exp(np.random.default_rng(1).uniform(np.log(0.1), np.log(10), mel.shape[1])).astype(np.float32)
    h = gmg.SSGMG(S.Mesh(S.SSHEX8, L, el, N), "elasticity", (adj, det), nodes, coeff=coeff, mu=mu, lam=lam,
                  coarse_rtol=1e-12)
    b = np.zeros((N, 3))
    b[:, 2] = -1
    b[nodes] = 0
    solver = S.cg(h.fine.A, h.mg, rtol=1e-10, max_it=300)
    x = zeros(3 * N)
    solver.apply(dev(b.ravel()), x)
    res = true_residual(h.fine.A, dev(b.ravel()), x)
    print(f"elasticity MG-PCG: {solver.iterations} its, residual {res:.2e}")
    assert solver.last_status == 0 and res < 1e-8


def test_elasticity_mg_pcg_stencil_form_matches_element_form():
    """The same L=8 elasticity MG-PCG (clamped z = 0, body force) with the operators in element form
    and in stencil form: identical FP64 iteration counts and residual histories to rounding, and the
    FP32 stencil preconditioner inside FP64 flexible CG reaches the FP64 tolerance."""
    L = 8
    mel, mpts = O.box_mesh(3, 3, 3)
    el, pts = O.sshex8_refine(L, mel, mpts)
    adj, det = O.hex8_adj_det(mel, mpts)
    N = pts.shape[1]
    nodes = np.where(pts[2] < 1e-6)[0].astype(np.int32)
    mu, lam = O.lame(1.0, 0.3)
    coeff = np.exp(np.random.default_rng(2).uniform(np.log(0.1), np.log(10), mel.shape[1])).astype(np.float32)
    b = np.zeros((N, 3))
    b[:, 2] = -1
    b[nodes] = 0
    runs = {}
    for form in (1, 2):
        S.set_option("elasticity_form", form)
        try:
            mesh = S.Mesh(S.SSHEX8, L, el, N)
            h = gmg.SSGMG(mesh, "elasticity", (adj, det), nodes, coeff=coeff, mu=mu, lam=lam, coarse_rtol=1e-12)
            solver = S.cg(h.fine.A, h.mg, rtol=1e-10, max_it=300)
            x = zeros(3 * N)
            solver.apply(dev(b.ravel()), x)
            res = true_residual(h.fine.A, dev(b.ravel()), x)
            h32 = gmg.SSGMG(mesh, "elasticity", (adj, det), nodes, coeff=coeff, mu=mu, lam=lam,
                            real_type=S.FLOAT32)
            fs = S.cg(h.fine.A, S.precision_bridge(h32.mg), flexible=True, rtol=1e-10, max_it=300)
            x32 = zeros(3 * N)
            fs.apply(dev(b.ravel()), x32)
            res32 = true_residual(h.fine.A, dev(b.ravel()), x32)
            runs[form] = (solver.iterations, res, host(x), fs.iterations, res32)
            print(f"form {form}: FP64 {solver.iterations} its (res {res:.2e}), mixed {fs.iterations} its "
                  f"(res {res32:.2e})")
            # the reference's PCG stops on the preconditioned residual: the true one lands near 1e-9
            assert solver.last_status == 0 and res < 1e-8
            assert fs.last_status == 0 and res32 < 1e-9
        finally:
            S.set_option("elasticity_form", 0)
    assert runs[1][0] == runs[2][0]
    assert np.linalg.norm(runs[2][2] - runs[1][2]) / np.linalg.norm(runs[1][2]) < 1e-9
    assert abs(runs[1][3] - runs[2][3]) <= 1


@pytest.mark.parametrize("real_type", [S.FLOAT64, S.FLOAT32])
def test_structured_extension_below_macro_mesh(real_type):
    """Hierarchy continued below the macro box (super-mesh levels): still converges, and the
    coarsest FP64 problem shrinks to the super mesh's corners."""
    L, n = 4, 8
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, n, coeff_seed=11)
    mesh = S.Mesh(S.SSHEX8, L, el, N)
    h64 = gmg.SSGMG(mesh, "laplacian", fff, nodes, coeff=coeff, macro_box=(n, 1.0))
    assert [lv.mesh.level for lv in h64.levels] == [4, 2, 1, 4, 2, 1]
    assert h64.levels[-1].mesh.nnodes == 8
    b = np.random.default_rng(0).uniform(0, 1, N)
    b[nodes] = 0
    if real_type == S.FLOAT64:
        solver = S.cg(h64.fine.A, h64.mg, rtol=1e-10, max_it=300)
    else:
        h32 = gmg.SSGMG(mesh, "laplacian", fff, nodes, coeff=coeff, real_type=S.FLOAT32, macro_box=(n, 1.0))
        solver = S.cg(h64.fine.A, S.precision_bridge(h32.mg), flexible=True, rtol=1e-10, max_it=300)
    x = zeros(N)
    solver.apply(dev(b), x)
    res = true_residual(h64.fine.A, dev(b), x)
    print(f"extended hierarchy ({'FP64' if real_type == S.FLOAT64 else 'FP32'}): {solver.iterations} its, "
          f"residual {res:.2e}")
    assert solver.last_status == 0 and res < 1e-8


def test_hex8_elasticity_mg_pcg_fp64_and_mixed_match_oracle():
    """hex8 linear elasticity on a 16^3 box (the 128^3 workload's problem at reduced size: E = 1,
    nu = 0.3, clamped at z = 0, body force (0, 0, -1)).  Geometric MG below the hex8 grid (super-mesh
    levels 8, 4, 2, 1), FP64 MG-PCG and the FP32 hierarchy (FP32 hex8 operator with bf16 geometry)
    inside FP64 flexible CG; both solutions against the oracle's CG solve of the same system."""
    n = 16
    el, pts = O.box_mesh(n + 1, n + 1, n + 1)
    adj, det = O.hex8_adj_det(el, pts)
    N = pts.shape[1]
    E = el.shape[1]
    nodes = np.where(pts[2] < 1e-6)[0].astype(np.int32)
    mu, lam = O.lame(1.0, 0.3)
    mesh = S.Mesh(S.HEX8, 1, el, N)
    h64 = gmg.SSGMG(mesh, "elasticity", (adj, det), nodes, mu=mu, lam=lam, macro_box=(n, 1.0))
    h32 = gmg.SSGMG(mesh, "elasticity", (adj, det), nodes, mu=mu, lam=lam, macro_box=(n, 1.0),
                    real_type=S.FLOAT32)
    assert [lv.mesh.level for lv in h64.levels] == [1, 8, 4, 2, 1]
    assert h32.levels[0].op.real_type == S.FLOAT32 and h32.levels[-1].op.real_type == S.FLOAT64
    b = np.zeros((N, 3))
    b[:, 2] = -1
    b[nodes] = 0
    b = b.ravel()
    dofs = (nodes[:, None].astype(np.int64) * 3 + np.arange(3)).ravel()
    A_o = oracle_constrained(lambda x, y: O.lib.oracle_hex8_elasticity_apply(E, el, adj, det, mu, lam, None, x, y),
                             dofs)
    d = np.zeros(3 * N)
    O.lib.oracle_hex8_elasticity_diag(E, el, adj, det, mu, lam, None, d)
    d[dofs] = 1
    xo, it_o, st_o, _ = SO.cg(A_o, b, np.zeros_like(b), rtol=1e-13, max_it=5000,
                              precond=lambda r, z: z.__iadd__(r / d))
    assert st_o == 0
    out = {}
    for name, P, flexible in (("FP64", h64.mg, False), ("mixed", S.precision_bridge(h32.mg), True)):
        solver = S.cg(h64.fine.A, P, flexible=flexible, rtol=1e-10, max_it=300)
        x = zeros(3 * N)
        solver.apply(dev(b), x)
        res = true_residual(h64.fine.A, dev(b), x)
        diff = np.linalg.norm(host(x) - xo) / np.linalg.norm(xo)
        out[name] = solver.iterations
        print(f"hex8 elasticity {name} MG-PCG: {solver.iterations} its, residual {res:.2e}, "
              f"difference to the oracle solve {diff:.2e}")
        assert solver.last_status == 0 and res < (1e-8 if name == "FP64" else 1e-10)
        assert diff < 1e-7
    assert out["mixed"] <= out["FP64"] + 3


def test_structured_extension_elasticity():
    L, n = 4, 4
    mel, mpts = O.box_mesh(n + 1, n + 1, n + 1)
    el, pts = O.sshex8_refine(L, mel, mpts)
    adj, det = O.hex8_adj_det(mel, mpts)
    N = pts.shape[1]
    nodes = np.where(pts[2] < 1e-6)[0].astype(np.int32)
    mu, lam = O.lame(1.0, 0.3)
    h = gmg.SSGMG(S.Mesh(S.SSHEX8, L, el, N), "elasticity", (adj, det), nodes, mu=mu, lam=lam,
                  macro_box=(n, 1.0))
    b = np.zeros((N, 3))
    b[:, 2] = -1
    b[nodes] = 0
    solver = S.cg(h.fine.A, h.mg, rtol=1e-10, max_it=300)
    x = zeros(3 * N)
    solver.apply(dev(b.ravel()), x)
    res = true_residual(h.fine.A, dev(b.ravel()), x)
    print(f"extended elasticity hierarchy: {solver.iterations} its, residual {res:.2e}")
    assert solver.last_status == 0 and res < 1e-8


@pytest.mark.parametrize("real_type", [0, 1])
def test_graph_replayed_vcycle_matches_eager(real_type):
    """The V-cycle (dense coarse solve) replayed as a HIP graph equals the eager V-cycle on every
    replay, with eager work between replays; PCG with the graph preconditioner converges in the
    same number of iterations."""
    L = 4
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 4, coeff_seed=3)
    mesh = S.Mesh(S.SSHEX8, L, el, N)
    h = gmg.SSGMG(mesh, "laplacian", fff, nodes, coeff=coeff, real_type=real_type, coarse="dense",
                  macro_box=(4, 1.0))
    P_eager = h.preconditioner(graph=False)
    P_graph = h.preconditioner(graph=True)
    rng = np.random.default_rng(5)
    z_graph = zeros(N)
    r = dev(rng.uniform(-1, 1, N))
    for k in range(4):
        r.copy_from(rng.uniform(-1, 1, N))
        z_eager = zeros(N)
        P_eager.apply(r, z_eager)
        z_graph.zero()
        P_graph.apply(r, z_graph)
        S.synchronize()
        err = np.linalg.norm(host(z_graph) - host(z_eager)) / np.linalg.norm(host(z_eager))
        assert err < (1e-12 if real_type == 0 else 1e-6), (k, err)
    captured, replays = S.graph_op_stats(P_graph)
    assert captured == 1 and replays == 3
    A64 = gmg.SSGMG(mesh, "laplacian", fff, nodes, coeff=coeff).fine.A if real_type else h.fine.A
    b = rng.uniform(0, 1, N)
    b[nodes] = 0
    its = []
    for P in (P_eager, P_graph):
        solver = S.cg(A64, P, flexible=True, rtol=1e-10, max_it=200)
        x = zeros(N)
        solver.apply(dev(b), x)
        rr = zeros(N)
        A64.apply(x, rr)
        assert np.linalg.norm(b - host(rr)) / np.linalg.norm(b) < 1e-10
        its.append(solver.iterations)
    assert its[0] == its[1]
    # the CG applies its preconditioner in zero-initial-guess mode: a second graph, replayed (around
    # the bridged FP32 V-cycle the flexible CG keeps z in FP32 and replays a graph of the bridge's
    # FP32 part of its own; with the option mixed_cg 0 it replays this one)
    if real_type:
        S.set_option("mixed_cg", 0)
        try:
            solver = S.cg(A64, P_graph, flexible=True, rtol=1e-10, max_it=200)
        finally:
            S.set_option("mixed_cg", 1)
        x = zeros(N)
        solver.apply(dev(b), x)
        assert solver.iterations == its[1]
    captured, replays2 = S.graph_op_stats(P_graph)
    assert captured == 2 and replays2 >= replays + its[1] - 2


def test_block_jacobi_smoother_parity_and_default_gmg():
    """Block-Jacobi stationary smoother (3x3 nodal blocks, omega 1/3, constrained rows as unit
    rows) against the oracle's, and MG-PCG with the reference's default smoothers on elasticity."""
    L = 2
    mel, mpts = O.box_mesh(3, 3, 3)
    el, pts = O.sshex8_refine(L, mel, mpts)
    adj, det = O.hex8_adj_det(mel, mpts)
    N = pts.shape[1]
    E = mel.shape[1]
    nodes = np.where(pts[2] < 1e-6)[0].astype(np.int32)
    mu, lam = O.lame(1.0, 0.3)
    op = S.linear_elasticity(S.Mesh(S.SSHEX8, L, el, N), adj, det, mu, lam)
    A = S.constrained(op, nodes, 3)
    A_o = oracle_constrained(
        lambda x, y: O.lib.oracle_sshex8_elasticity_apply(L, E, el, adj, det, mu, lam, None, x, y),
        (nodes[:, None] * 3 + np.arange(3)).ravel().astype(np.int32))
    sym = np.zeros(6 * N)
    O.lib.oracle_sshex8_elasticity_block_diag(L, E, el, adj, det, mu, lam, None, sym)
    mask = np.zeros(3 * N, np.uint8)
    mask[(nodes[:, None] * 3 + np.arange(3)).ravel()] = 1
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, 3 * N)
    x0 = rng.uniform(-1, 1, 3 * N)
    b[mask == 1] = 0
    x0[mask == 1] = 0
    xo = SO.BlockJacobi(A_o, sym, mask, omega=1.0 / 3, steps=3)(b, x0.copy())
    d6 = zeros(6 * N)
    op.block_diag_sym(d6)
    sm = S.block_jacobi(A, d6, dev(mask, np.uint8), omega=1.0 / 3, steps=3)
    x = dev(x0)
    sm.apply(dev(b), x)
    S.synchronize()
    assert np.linalg.norm(host(x) - xo) / np.linalg.norm(xo) < 1e-12
    # reference default GMG (block-Jacobi smoothing) as a preconditioner
    L2 = 4
    el2, pts2 = O.sshex8_refine(L2, mel, mpts)
    N2 = pts2.shape[1]
    nodes2 = np.where(pts2[2] < 1e-6)[0].astype(np.int32)
    h = gmg.SSGMG(S.Mesh(S.SSHEX8, L2, el2, N2), "elasticity", (adj, det), nodes2, mu=mu, lam=lam,
                  smoother="jacobi", coarse_rtol=1e-12)
    rhs = np.zeros((N2, 3))
    rhs[:, 2] = -1
    rhs[nodes2] = 0
    solver = S.cg(h.fine.A, h.mg, rtol=1e-10, max_it=400)
    xs = zeros(3 * N2)
    solver.apply(dev(rhs.ravel()), xs)
    res = true_residual(h.fine.A, dev(rhs.ravel()), xs)
    print(f"elasticity MG-PCG with block-Jacobi smoothing: {solver.iterations} its, residual {res:.2e}")
    assert solver.last_status == 0 and res < 1e-8


@pytest.mark.parametrize("L,rt", [(4, S.FLOAT32), (8, S.FLOAT64), (8, S.FLOAT32), (16, S.FLOAT64), (16, S.FLOAT32)])
@pytest.mark.parametrize("jacobi", [False, True])
def test_fused_chebyshev_steps_match_separate_kernels(L, rt, jacobi):
    """Chebyshev3 on a constrained sshex8 elasticity operator: the steps run inside the matrix-core
    apply (interior nodes) and the skeleton gather (skeleton nodes, Dirichlet rows) must give the
    words of the separate product + update kernels, for the general-guess start and the updates."""
    dt = np.float64 if rt == S.FLOAT64 else np.float32
    el_m, pts_m = O.box_mesh(3, 3, 3)
    el, pts = O.sshex8_refine(L, el_m, pts_m)
    N = pts.shape[1]
    mesh = S.Mesh(S.SSHEX8, L, el, N)
    adj, det = O.hex8_adj_det(el_m, pts_m)
    mu, lam = O.lame(1.0, 0.3)
    coeff = np.exp(np.random.default_rng(L).uniform(-1, 1, el_m.shape[1])).astype(np.float32)
    op = S.linear_elasticity(mesh, adj, det, mu, lam, coeff, real_type=rt)
    nodes = np.where(pts[2] < 1e-6)[0].astype(np.int32)
    A = S.constrained(op, nodes, 3)
    inv = None
    if jacobi:
        d = zeros(3 * N, dt)
        A.diag(d)
        inv = dev((1.0 / host(d)).astype(dt), dt)
    init = np.ones(3 * N, dt)
    init[(nodes[:, None] * 3 + np.arange(3)).ravel()] = 0
    sm = S.cheb3(A, max_it=3, scale_eig_max=1.02, init_vector=dev(init, dt), inv_diag=inv)
    rng = np.random.default_rng(1)
    rhs = rng.uniform(-1, 1, 3 * N).astype(dt)
    rhs[(nodes[:, None] * 3 + np.arange(3)).ravel()] = 0
    x0 = rng.uniform(-1, 1, 3 * N).astype(dt)
    out = {}
    try:
        for fused in (1, 0):
            S.set_option("fused_chebyshev", fused)
            x = dev(x0, dt)
            sm.apply(dev(rhs, dt), x)
            S.synchronize()
            out[fused] = host(x)
    finally:
        S.set_option("fused_chebyshev", 0)
    np.testing.assert_array_equal(out[1], out[0])
    assert np.linalg.norm(out[1] - x0) > 0


def test_flexible_cg_matches_oracle_with_fixed_jacobi_preconditioner():
    """The product's flexible PCG (mode 1) against the oracle's flexible_pcg on a variable-
    coefficient sshex8 Poisson problem with a fixed SPD (Jacobi) preconditioner: same iteration
    count, the same residual history to rounding, the same solution."""
    L = 4
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 3, coeff_seed=5)
    E = mel.shape[1]
    A_o = oracle_constrained(lambda x, y: O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, coeff, x, y), nodes)
    d = np.zeros(N)
    O.lib.oracle_sshex8_laplacian_diag(L, E, el, fff, coeff, d)
    d[nodes] = 1

    def jac(r, z):
        z += r / d
    b = np.random.default_rng(7).uniform(0, 1, N)
    b[nodes] = 0
    xo, it_o, st_o, hist_o = SO.flexible_pcg(A_o, b, np.zeros(N), jac, rtol=1e-10)
    assert st_o == 0
    A = S.constrained(S.laplacian(S.Mesh(S.SSHEX8, L, el, N), fff, coeff), nodes)
    M = S.inverse_diagonal(dev(d))
    solver = S.cg(A, M, flexible=True, rtol=1e-10)
    x = zeros(N)
    solver.apply(dev(b), x)
    hist = solver.history()
    print(f"flexible CG: gpu {solver.iterations} its, oracle {it_o} its")
    assert solver.last_status == 0 and solver.iterations == it_o
    m = min(len(hist), len(hist_o))
    # relative residuals agree to 1e-7 of their value, or to 1e-13 absolute near the 1e-10 stop
    np.testing.assert_allclose(hist[:m], hist_o[:m], rtol=1e-7, atol=1e-13)
    assert np.linalg.norm(host(x) - xo) / np.linalg.norm(xo) < 1e-9
    assert true_residual(A, dev(b), x) < 1e-10


@pytest.mark.parametrize("omega,steps", [(1.0, 3), (2.0 / 3.0, 5)])
def test_scalar_jacobi_stationary_iteration_matches_oracle(omega, steps):
    """StationaryIteration with scaled Jacobi (the reference's Laplacian GMG smoother, omega = 1,
    three sweeps) against oracle/solvers.Jacobi on the constrained sshex8 Poisson operator."""
    L = 4
    mel, mpts, el, pts, fff, N, nodes, coeff = _ss_poisson(L, 3, coeff_seed=6)
    E = mel.shape[1]
    A_o = oracle_constrained(lambda x, y: O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, coeff, x, y), nodes)
    A = S.constrained(S.laplacian(S.Mesh(S.SSHEX8, L, el, N), fff, coeff), nodes)
    dd = zeros(N)
    A.diag(dd)
    d = host(dd)
    rng = np.random.default_rng(8)
    b = rng.uniform(-1, 1, N)
    b[nodes] = 0
    x0 = rng.uniform(-1, 1, N)
    x0[nodes] = 0
    xo = SO.Jacobi(A_o, d, omega=omega, steps=steps)(b, x0.copy())
    sm = S.jacobi(A, dd, omega=omega, steps=steps)
    x = dev(x0)
    sm.apply(dev(b), x)
    S.synchronize()
    err = np.linalg.norm(host(x) - xo) / np.linalg.norm(xo)
    print(f"Jacobi omega={omega:.3f} steps={steps}: relative difference {err:.2e}")
    assert err < 1e-12
