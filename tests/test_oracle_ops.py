"""Pins the oracle's geometry, element matrices and operators.

Checks are chosen so a plausible mistake (a dropped term, a transposed operand,
a wrong index) fails: closed forms, an independent numpy quadrature assembly,
the reference's stencil known-answer test, null spaces, Galerkin identities.
"""
import numpy as np
import pytest

import oracle_lib as O

RNG = np.random.default_rng(1234)
GAUSS3 = (np.array([0.5 - 0.5 * np.sqrt(0.6), 0.5, 0.5 + 0.5 * np.sqrt(0.6)]), np.array([5, 8, 5]) / 18.0)


def q1_ref_grads(q):
    """Reference-cube gradients of the 8 trilinear basis functions at q (8x3)."""
    g = np.zeros((8, 3))
    for n, c in enumerate(O.HEX8_CORNERS):
        for d in range(3):
            r = 1.0
            for k in range(3):
                if k == d:
                    r *= 1.0 if c[k] else -1.0
                else:
                    r *= q[k] if c[k] else 1 - q[k]
            g[n, d] = r
    return g


def numpy_laplacian_matrix(G):
    """K_ij = int grad_i . G grad_j over the unit cube, 3-point Gauss (exact)."""
    x, w = GAUSS3
    K = np.zeros((8, 8))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                g = q1_ref_grads((x[i], x[j], x[k]))
                K += w[i] * w[j] * w[k] * g @ G @ g.T
    return K


def numpy_elasticity_matrix(mu, lam, J):
    """Displacement-based Q1 elasticity with a constant Jacobian J, 3-point Gauss,
    index = component*8 + node.  Built from B (strain) matrices in Voigt form."""
    x, w = GAUSS3
    Jinv = np.linalg.inv(J)
    det = np.linalg.det(J)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[:3, :3] += 2 * mu * np.eye(3)
    D[3:, 3:] = mu * np.eye(3)
    K = np.zeros((24, 24))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                g = q1_ref_grads((x[i], x[j], x[k])) @ Jinv  # physical gradients (8x3)
                B = np.zeros((6, 24))
                for n in range(8):
                    B[0, 0 * 8 + n] = g[n, 0]
                    B[1, 1 * 8 + n] = g[n, 1]
                    B[2, 2 * 8 + n] = g[n, 2]
                    B[3, 0 * 8 + n] = g[n, 1]; B[3, 1 * 8 + n] = g[n, 0]
                    B[4, 1 * 8 + n] = g[n, 2]; B[4, 2 * 8 + n] = g[n, 1]
                    B[5, 0 * 8 + n] = g[n, 2]; B[5, 2 * 8 + n] = g[n, 0]
                K += w[i] * w[j] * w[k] * det * B.T @ D @ B
    return K


def sym6(G):
    return np.array([G[0, 0], G[0, 1], G[0, 2], G[1, 1], G[1, 2], G[2, 2]])


def rand_affine_hex():
    """Corners of a random well-shaped parallelepiped (float32 words)."""
    J = np.eye(3) + 0.3 * RNG.uniform(-1, 1, (3, 3))
    o = RNG.uniform(-1, 1, 3)
    pts = np.array([o + J @ np.array(c, float) for c in O.HEX8_CORNERS]).T.astype(np.float32)
    return pts, J


# ---------------- geometry ----------------

def test_tet4_fff_closed_forms():
    el = np.arange(4, dtype=np.int32).reshape(4, 1)
    right = np.array([[0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32)
    np.testing.assert_allclose(O.tet4_fff(el, right)[0], np.array([1, 0, 0, 1, 0, 1]) / 6, rtol=1e-7)
    # J = 2I: J^-1 J^-T det/6 = I/3, while J J^T det/6 would be 16/3
    np.testing.assert_allclose(O.tet4_fff(el, 2 * right)[0], np.array([1, 0, 0, 1, 0, 1]) / 3, rtol=1e-7)


def test_tet4_fff_general():
    J = np.eye(3) + 0.4 * RNG.uniform(-1, 1, (3, 3))
    p0 = RNG.uniform(-1, 1, 3)
    pts = np.stack([p0, p0 + J[:, 0], p0 + J[:, 1], p0 + J[:, 2]], axis=1).astype(np.float32)
    el = np.arange(4, dtype=np.int32).reshape(4, 1)
    Jf = (pts[:, 1:].astype(np.float64) - pts[:, :1])
    Ji = np.linalg.inv(Jf)
    np.testing.assert_allclose(O.tet4_fff(el, pts)[0], sym6(Ji @ Ji.T * np.linalg.det(Jf) / 6), rtol=2e-6)


def test_tet4_fff_is_float32_evaluation():
    """tet4_fff (operators/tet4/tet4_inline_cpu.hpp:224-278) keeps its temporaries in float32: the
    words agree with J^-1 J^-T det/6 evaluated in float64 to float32 rounding error, and differ
    from that value rounded once to float32 in some words (a float64 restatement would match it)."""
    rng = np.random.default_rng(12)
    el, pts = O.box_mesh(7, 7, 7, tet4=True)
    pts = (pts + 0.04 * rng.uniform(-1, 1, pts.shape)).astype(np.float32)
    got = O.tet4_fff(el, pts)
    want = np.empty_like(got, dtype=np.float64)
    for e in range(el.shape[1]):
        v = pts[:, el[:, e]].astype(np.float64)
        J = v[:, 1:] - v[:, :1]
        Ji = np.linalg.inv(J)
        want[e] = sym6(Ji @ Ji.T * np.linalg.det(J) / 6)
    scale = np.abs(want).max(axis=1, keepdims=True)
    assert np.max(np.abs(got - want) / scale) < 2e-5
    assert np.count_nonzero(got != want.astype(np.float32)) > 0.05 * got.size


def test_hex8_fff_box_and_parallelepiped():
    el = np.arange(8, dtype=np.int32).reshape(8, 1)
    a, b, c = 2.0, 0.5, 4.0
    pts = np.array([[a * x, b * y, c * z] for (x, y, z) in O.HEX8_CORNERS]).T.astype(np.float32)
    np.testing.assert_allclose(O.hex8_fff(el, pts)[0], [b * c / a, 0, 0, a * c / b, 0, a * b / c], rtol=1e-7)
    pts, J = rand_affine_hex()
    Jf = np.array([[1.0]])  # recompute J from the float32 corners
    Jf = np.stack([pts[:, 1] - pts[:, 0], pts[:, 3] - pts[:, 0], pts[:, 4] - pts[:, 0]], axis=1).astype(np.float64)
    Ji = np.linalg.inv(Jf)
    np.testing.assert_allclose(O.hex8_fff(el, pts)[0], sym6(Ji @ Ji.T * np.linalg.det(Jf)), rtol=1e-5, atol=1e-6)
    adj, det = O.hex8_adj_det(el, pts)
    np.testing.assert_allclose(det[0], np.linalg.det(Jf), rtol=1e-6)
    np.testing.assert_allclose(adj[0].reshape(3, 3), Ji * np.linalg.det(Jf), rtol=1e-5, atol=1e-6)


# ---------------- element matrices ----------------

@pytest.mark.parametrize("seed", range(3))
def test_hex8_laplacian_matrix_vs_quadrature(seed):
    A = np.random.default_rng(seed).uniform(-1, 1, (3, 3))
    G = A @ A.T + np.eye(3)
    np.testing.assert_allclose(O.hex8_laplacian_matrix(sym6(G)), numpy_laplacian_matrix(G), rtol=1e-13, atol=1e-14)


def test_tet4_laplacian_matrix_vs_direct():
    J = np.eye(3) + 0.4 * RNG.uniform(-1, 1, (3, 3))
    Ji = np.linalg.inv(J)
    det = np.linalg.det(J)
    grads = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) @ Ji
    K = grads @ grads.T * det / 6
    np.testing.assert_allclose(O.tet4_laplacian_matrix(sym6(Ji @ Ji.T * det / 6)), K, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("seed", range(3))
def test_hex8_elasticity_matrix_vs_quadrature(seed):
    r = np.random.default_rng(seed)
    J = np.eye(3) + 0.3 * r.uniform(-1, 1, (3, 3))
    mu, lam = O.lame(r.uniform(0.5, 2), 0.3)
    det = np.linalg.det(J)
    adj = np.linalg.inv(J) * det
    Ko = O.hex8_elasticity_matrix(mu, lam, adj.ravel(), det)
    np.testing.assert_allclose(Ko, numpy_elasticity_matrix(mu, lam, J), rtol=1e-12, atol=1e-13)
    # rigid-body modes of the physical element are in the null space
    X = np.array([J @ np.array(c, float) for c in O.HEX8_CORNERS])
    modes = []
    for d in range(3):
        t = np.zeros(24); t[d * 8:(d + 1) * 8] = 1; modes.append(t)
    for (a, b) in ((0, 1), (1, 2), (0, 2)):
        m = np.zeros(24); m[a * 8:(a + 1) * 8] = -X[:, b]; m[b * 8:(b + 1) * 8] = X[:, a]; modes.append(m)
    for m in modes:
        assert np.linalg.norm(Ko @ m) < 1e-12 * np.linalg.norm(Ko)


# ---------------- operators ----------------

def test_stencil_known_answer():
    """operators/stencil/tests/sfem_StencilTest.cpp:143-216: with unit FFF the
    27-point Q1 stencil applied to x^2+y^2+z^2 is -6 on every interior node.
    Here: one macro element [0,L]^3 so that sub_fff = h*FFF = I."""
    L = 8
    pts = np.array([[L * x, L * y, L * z] for (x, y, z) in O.HEX8_CORNERS]).T.astype(np.float32)
    mel = np.arange(8, dtype=np.int32).reshape(8, 1)
    el, fpts = O.sshex8_refine(L, mel, pts)
    fff = O.hex8_fff(mel, pts)
    np.testing.assert_allclose(fff[0], [L, 0, 0, L, 0, L], rtol=1e-7)
    u = (fpts.astype(np.float64) ** 2).sum(0)
    y = np.zeros_like(u)
    O.lib.oracle_sshex8_laplacian_apply(L, 1, el, fff, None, u, y)
    interior = np.all((fpts > 0.5) & (fpts < L - 0.5), axis=0)
    np.testing.assert_allclose(y[interior], -6.0, rtol=0, atol=1e-10)


def _cube_problem(kind, n=3, L=4, skew=True):
    mel, mpts = O.box_mesh(n + 1, n + 1, n + 1, tet4=(kind == "tet4"))
    if skew:  # affine skew keeps macro elements parallelepipeds
        A = np.array([[1.0, 0.2, 0.1], [0.0, 1.1, 0.15], [0.05, 0.0, 0.9]])
        mpts = (A @ mpts.astype(np.float64)).astype(np.float32)
    return mel, mpts


def test_laplacian_null_space_and_symmetry():
    for kind in ("tet4", "hex8", "sshex8"):
        mel, mpts = _cube_problem(kind)
        if kind == "tet4":
            fff = O.tet4_fff(mel, mpts)
            n = mpts.shape[1]
            apply = lambda x, y: O.lib.oracle_tet4_laplacian_apply(mel.shape[1], mel, fff, x, y)
        elif kind == "hex8":
            fff = O.hex8_fff(mel, mpts)
            n = mpts.shape[1]
            apply = lambda x, y: O.lib.oracle_hex8_laplacian_apply(mel.shape[1], mel, fff, None, x, y)
        else:
            L = 4
            el, pts = O.sshex8_refine(L, mel, mpts)
            fff = O.hex8_fff(mel, mpts)
            n = pts.shape[1]
            apply = lambda x, y: O.lib.oracle_sshex8_laplacian_apply(L, mel.shape[1], el, fff, None, x, y)
        y = np.zeros(n)
        apply(np.ones(n), y)
        assert np.abs(y).max() < 1e-12
        a, b = RNG.uniform(-1, 1, n), RNG.uniform(-1, 1, n)
        ya, yb = np.zeros(n), np.zeros(n)
        apply(a, ya)
        apply(b, yb)
        assert abs(ya @ b - yb @ a) < 1e-12 * abs(ya @ b)
        assert a @ ya > 0


def _micro_hex8(L, el):
    """Micro-element hex8 connectivity of an sshex8 mesh (E*L^3 elements)."""
    rows = []
    for zi in range(L):
        for yi in range(L):
            for xi in range(L):
                rows.append([O.lidx(L, xi + cx, yi + cy, zi + cz) for (cx, cy, cz) in O.HEX8_CORNERS])
    rows = np.array(rows)  # [L^3][8]
    return np.ascontiguousarray(el[rows.T].reshape(8, -1))  # [8][L^3*E], element-major within row


def test_sshex8_laplacian_vs_assembled_micro_mesh():
    """sshex8 apply == hex8 apply on the explicit micro-element mesh whose FFF
    is h * macro FFF (brute force over the refined mesh)."""
    L = 4
    mel, mpts = _cube_problem("hex8", n=2)
    el, pts = O.sshex8_refine(L, mel, mpts)
    fff = O.hex8_fff(mel, mpts)
    coeff = RNG.uniform(0.5, 2.0, mel.shape[1]).astype(np.float32)
    micro = _micro_hex8(L, el)
    E = mel.shape[1]
    # micro element k*E + e belongs to macro e
    mfff = np.repeat(fff[None, :, :], L ** 3, axis=0).reshape(-1, 6)
    mcoef = np.tile(coeff * np.float32(1.0 / L), L ** 3).astype(np.float32)
    n = pts.shape[1]
    x = RNG.uniform(-1, 1, n)
    y1, y2 = np.zeros(n), np.zeros(n)
    O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, coeff, x, y1)
    O.lib.oracle_hex8_laplacian_apply(micro.shape[1], micro, mfff, mcoef, x, y2)
    np.testing.assert_allclose(y1, y2, rtol=1e-12, atol=1e-12 * np.abs(y2).max())
    d1, d2 = np.zeros(n), np.zeros(n)
    O.lib.oracle_sshex8_laplacian_diag(L, E, el, fff, coeff, d1)
    O.lib.oracle_hex8_laplacian_diag(micro.shape[1], micro, mfff, mcoef, d2)
    np.testing.assert_allclose(d1, d2, rtol=1e-12)


def test_sshex8_elasticity_vs_assembled_micro_mesh():
    L = 2
    mel, mpts = _cube_problem("hex8", n=2)
    el, pts = O.sshex8_refine(L, mel, mpts)
    adj, det = O.hex8_adj_det(mel, mpts)
    mu, lam = O.lame(1.0, 0.3)
    E = mel.shape[1]
    coeff = RNG.uniform(0.1, 10.0, E).astype(np.float32)
    micro = _micro_hex8(L, el)
    h = np.float32(1.0 / L)
    madj = np.repeat(adj[None] * h * h, L ** 3, axis=0).reshape(-1, 9).astype(np.float32)
    mdet = np.repeat(det[None] * h * h * h, L ** 3, axis=0).reshape(-1).astype(np.float32)
    mcoef = np.tile(coeff, L ** 3).astype(np.float32)
    n = pts.shape[1] * 3
    x = RNG.uniform(-1, 1, n)
    y1, y2 = np.zeros(n), np.zeros(n)
    O.lib.oracle_sshex8_elasticity_apply(L, E, el, adj, det, mu, lam, coeff, x, y1)
    O.lib.oracle_hex8_elasticity_apply(micro.shape[1], micro, madj, mdet, mu, lam, mcoef, x, y2)
    np.testing.assert_allclose(y1, y2, rtol=1e-11, atol=1e-11 * np.abs(y2).max())
    d1, d2 = np.zeros(n), np.zeros(n)
    O.lib.oracle_sshex8_elasticity_diag(L, E, el, adj, det, mu, lam, coeff, d1)
    O.lib.oracle_hex8_elasticity_diag(micro.shape[1], micro, madj, mdet, mu, lam, mcoef, d2)
    np.testing.assert_allclose(d1, d2, rtol=1e-11)
    bd = np.zeros(pts.shape[1] * 6)
    O.lib.oracle_sshex8_elasticity_block_diag(L, E, el, adj, det, mu, lam, coeff, bd)
    np.testing.assert_allclose(bd.reshape(-1, 6)[:, [0, 3, 5]].ravel(), d1, rtol=1e-14)


def test_elasticity_rigid_body_modes_global():
    L = 4
    mel, mpts = _cube_problem("hex8", n=2)
    el, pts = O.sshex8_refine(L, mel, mpts)
    adj, det = O.hex8_adj_det(mel, mpts)
    mu, lam = O.lame(1.0, 0.3)
    X = pts.astype(np.float64)
    n = pts.shape[1]
    for m in range(6):
        u = np.zeros((n, 3))
        if m < 3:
            u[:, m] = 1
        else:
            a, b = [(0, 1), (1, 2), (0, 2)][m - 3]
            u[:, a] = -X[b]
            u[:, b] = X[a]
        y = np.zeros(3 * n)
        O.lib.oracle_sshex8_elasticity_apply(L, mel.shape[1], el, adj, det, mu, lam, None, u.ravel(), y)
        assert np.abs(y).max() < 2e-6  # float32 geometry words limit the cancellation


# ---------------- transfers ----------------

def _ss_levels(n, L, skew=True):
    mel, mpts = _cube_problem("hex8", n=n, skew=skew)
    el, pts = O.sshex8_refine(L, mel, mpts)
    return mel, mpts, el, pts


def test_prolongation_reproduces_linear_fields():
    mel, mpts, el, pts = _ss_levels(2, 8)
    for Lc in (4, 2):
        ec = O.derefine(8, Lc, el)
        nc = ec.max() + 1
        a = np.array([0.3, -1.2, 0.7])
        coarse = a @ pts[:, :nc].astype(np.float64) + 0.25
        fine = np.zeros(pts.shape[1])
        O.lib.oracle_sshex8_prolongate(mel.shape[1], Lc, ec, 8, el, 1, coarse, fine)
        np.testing.assert_allclose(fine, a @ pts.astype(np.float64) + 0.25, atol=2e-6)
    # hex8 level 1 -> sshex8 level 2
    e2 = O.derefine(8, 2, el)
    n2 = e2.max() + 1
    n1 = mpts.shape[1]
    fine = np.zeros(n2)
    O.lib.oracle_sshex8_hierarchical_prolongation(2, mel.shape[1], e2, 1, np.ones(n1), fine)
    np.testing.assert_allclose(fine, 1.0, rtol=1e-15)


def _dense(apply, n_in, n_out):
    A = np.zeros((n_out, n_in))
    for j in range(n_in):
        e = np.zeros(n_in); e[j] = 1
        y = np.zeros(n_out)
        apply(e, y)
        A[:, j] = y
    return A


def test_restriction_is_prolongation_transpose_and_galerkin():
    L = 4
    mel, mpts, el, pts = _ss_levels(2, L)
    E = mel.shape[1]
    nf = pts.shape[1]
    count = O.incidence(el, nf)
    fff = O.hex8_fff(mel, mpts)
    ec = O.derefine(L, 2, el)
    nc = ec.max() + 1
    P = _dense(lambda c, f: O.lib.oracle_sshex8_prolongate(E, 2, ec, L, el, 1, c, f), nc, nf)
    R = _dense(lambda f, c: O.lib.oracle_sshex8_restrict(E, L, el, count, 2, ec, 1, f, c), nf, nc)
    np.testing.assert_allclose(R, P.T, atol=1e-14)
    Af = _dense(lambda x, y: O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, None, x, y), nf, nf)
    Ac = _dense(lambda x, y: O.lib.oracle_sshex8_laplacian_apply(2, E, ec, fff, None, x, y), nc, nc)
    # nested Q1 spaces with exact integration: P^T A_f P == A_c
    np.testing.assert_allclose(P.T @ Af @ P, Ac, atol=1e-12 * np.abs(Ac).max())
    # hex8 (level 1) <-> sshex8 level 2
    h8 = O.to_hex8(1, O.derefine(L, 1, el))
    n1 = mpts.shape[1]
    c2 = O.incidence(ec, nc)
    P1 = _dense(lambda c, f: O.lib.oracle_sshex8_hierarchical_prolongation(2, E, ec, 1, c, f), n1, nc)
    R1 = _dense(lambda f, c: O.lib.oracle_sshex8_hierarchical_restriction(2, E, ec, c2, 1, f, c), nc, n1)
    np.testing.assert_allclose(R1, P1.T, atol=1e-14)
    A1 = _dense(lambda x, y: O.lib.oracle_hex8_laplacian_apply(E, h8, fff, None, x, y), n1, n1)
    np.testing.assert_allclose(P1.T @ Ac @ P1, A1, atol=1e-12 * np.abs(A1).max())


def test_vector_transfers_act_per_component():
    L = 4
    mel, mpts, el, pts = _ss_levels(1, L)
    E = mel.shape[1]
    ec = O.derefine(L, 2, el)
    nc, nf = ec.max() + 1, pts.shape[1]
    c = RNG.uniform(-1, 1, (nc, 3))
    f3 = np.zeros((nf, 3))
    O.lib.oracle_sshex8_prolongate(E, 2, ec, L, el, 3, np.ascontiguousarray(c.ravel()), f3.reshape(-1))
    for d in range(3):
        f1 = np.zeros(nf)
        O.lib.oracle_sshex8_prolongate(E, 2, ec, L, el, 1, np.ascontiguousarray(c[:, d]), f1)
        np.testing.assert_allclose(f3[:, d], f1, rtol=1e-15)


def test_dirichlet_helpers():
    x = np.arange(12, dtype=np.float64)
    nodes = np.array([1, 3], np.int32)
    O.lib.oracle_constraint_nodes_to_value_vec(2, nodes, 3, 1, -5.0, x)
    assert x[4] == -5 and x[10] == -5 and x[3] == 3
    y = np.zeros(12)
    O.lib.oracle_constraint_nodes_copy_vec(2, nodes, 3, 2, x, y)
    assert y[5] == 5 and y[11] == 11 and y.sum() == 16


def _dense_micro_assembly(L, el, E, K, bs):
    """Dense global matrix of the stored-matrix (ss:em) operator, assembled micro element by
    micro element from the explicit corner lists (independent of the oracle's loops)."""
    micro = _micro_hex8(L, el)  # [8][L^3 * E], macro e owns columns k*E + e
    n = (el.max() + 1) * bs
    A = np.zeros((n, n))
    for col in range(micro.shape[1]):
        e = col % E
        nodes = micro[:, col]
        dofs = np.array([nodes[v] * bs + c for c in range(bs) for v in range(8)])  # component*8 + node
        A[np.ix_(dofs, dofs)] += K[e].reshape(8 * bs, 8 * bs)
    return A


@pytest.mark.parametrize("bs", [1, 3])
def test_em_apply_vs_dense_assembly(bs):
    """Random (non-symmetric) stored matrices: the oracle's ss:em apply equals the densely
    assembled operator, so the corner order and the component*8 + node packing are pinned."""
    L = 2
    mel, mpts = _cube_problem("hex8", n=2)
    el, pts = O.sshex8_refine(L, mel, mpts)
    E = mel.shape[1]
    K = RNG.normal(size=(E, (8 * bs) ** 2))
    n = pts.shape[1] * bs
    x = RNG.uniform(-1, 1, n)
    y = np.zeros(n)
    O.lib.oracle_sshex8_em_apply(L, E, el, bs, np.ascontiguousarray(K), x, y)
    A = _dense_micro_assembly(L, el, E, K, bs)
    np.testing.assert_allclose(y, A @ x, rtol=1e-12, atol=1e-12 * np.abs(A @ x).max())


def test_em_apply_with_geometry_matrices_equals_operators():
    """ss:em with the matrices sshex8_laplacian_element_matrix builds equals the FFF Laplacian;
    with the elasticity sub-element matrices it equals the adjugate elasticity operator."""
    L = 4
    mel, mpts = _cube_problem("hex8", n=2)
    el, pts = O.sshex8_refine(L, mel, mpts)
    E = mel.shape[1]
    fff = O.hex8_fff(mel, mpts)
    coeff = RNG.uniform(0.5, 2.0, E).astype(np.float32)
    K = np.zeros((E, 64))
    O.lib.oracle_sshex8_laplacian_element_matrices(L, E, fff, coeff, K)
    n = pts.shape[1]
    x = RNG.uniform(-1, 1, n)
    y1, y2 = np.zeros(n), np.zeros(n)
    O.lib.oracle_sshex8_em_apply(L, E, el, 1, K, x, y1)
    O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, coeff, x, y2)
    np.testing.assert_allclose(y1, y2, rtol=1e-12, atol=1e-13 * np.abs(y2).max())
    adj, det = O.hex8_adj_det(mel, mpts)
    mu, lam = O.lame(1.0, 0.3)
    K3 = np.zeros((E, 576))
    O.lib.oracle_sshex8_elasticity_element_matrices(L, E, adj, det, mu, lam, coeff, K3)
    x3 = RNG.uniform(-1, 1, 3 * n)
    y1, y2 = np.zeros(3 * n), np.zeros(3 * n)
    O.lib.oracle_sshex8_em_apply(L, E, el, 3, K3, x3, y1)
    O.lib.oracle_sshex8_elasticity_apply(L, E, el, adj, det, mu, lam, coeff, x3, y2)
    np.testing.assert_allclose(y1, y2, rtol=1e-12, atol=1e-13 * np.abs(y2).max())


@pytest.mark.parametrize("L", [2, 3, 8])
def test_sshex8_laplacian_stencil_path_equals_element_path(L):
    """The stencil + surface-stencil evaluation (the reference's timed path,
    affine_sshex8_laplacian_stencil_apply_fff) gives the element-by-element operator; with unit
    FFF on [0,L]^3 it reproduces the -6 known answer of operators/stencil/tests/sfem_StencilTest.cpp."""
    mel, mpts = _cube_problem("hex8", n=2)
    el, pts = O.sshex8_refine(L, mel, mpts)
    fff = O.hex8_fff(mel, mpts)
    coeff = RNG.uniform(0.5, 2.0, mel.shape[1]).astype(np.float32)
    x = RNG.uniform(-1, 1, pts.shape[1])
    y1, y2 = np.zeros_like(x), np.full_like(x, 0.25)
    O.lib.oracle_sshex8_laplacian_apply(L, mel.shape[1], el, fff, coeff, x, y1)
    O.lib.oracle_sshex8_laplacian_stencil_apply(L, mel.shape[1], el, fff, coeff, x, y2)
    np.testing.assert_allclose(y2 - 0.25, y1, rtol=1e-12, atol=1e-13 * np.abs(y1).max())
    pts1 = np.array([[L * a, L * b, L * c] for (a, b, c) in O.HEX8_CORNERS]).T.astype(np.float32)
    el1, fp1 = O.sshex8_refine(L, np.arange(8, dtype=np.int32).reshape(8, 1), pts1)
    u = (fp1.astype(np.float64) ** 2).sum(0)
    y = np.zeros_like(u)
    O.lib.oracle_sshex8_laplacian_stencil_apply(L, 1, el1, O.hex8_fff(np.arange(8, dtype=np.int32).reshape(8, 1), pts1),
                                                None, u, y)
    interior = np.all((fp1 > 0.5) & (fp1 < L - 0.5), axis=0)
    if interior.any():
        np.testing.assert_allclose(y[interior], -6.0, rtol=0, atol=1e-10)
