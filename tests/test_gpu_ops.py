"""GPU parity of the HIP operators against the CPU oracle, through the C ABI.

Bar: FP64 apply/diag within 1e-12 relative (2-norm) of the oracle on the same float32
geometry words; transfers likewise; integer data (connectivity after upload/derefine)
bit-exact.  FP32/BF16 operators state their error against the FP64 oracle (< 1e-2).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

import sfem_amd as S  # noqa: E402

RTOL = 1e-12


def dev(a, dtype=np.float64):
    return S.DeviceArray.from_numpy(np.asarray(a), dtype)


def host(t):
    return t.numpy()


def zeros(n, dtype=np.float64):
    return S.DeviceArray(n, dtype)


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def skewed_box(n, tet4=False, seed=0, perturb=0.0):
    el, pts = O.box_mesh(n + 1, n + 1, n + 1, tet4=tet4)
    A = np.array([[1.0, 0.2, 0.1], [0.0, 1.1, 0.15], [0.05, 0.0, 0.9]])
    pts = (A @ pts.astype(np.float64)).astype(np.float32)
    if perturb:
        pts = (pts + perturb * np.random.default_rng(seed).uniform(-1, 1, pts.shape)).astype(np.float32)
    return el, pts


def gpu_apply(op, x):
    y = zeros(op.rows, x.dtype)
    op.apply(x, y)
    S.synchronize()
    return y


def test_device_present():
    assert S.device_count() >= 1


# ---------------------------------------------------------------- Laplacians
def test_tet4_laplacian_parity():
    el, pts = skewed_box(8, tet4=True, perturb=0.02)
    fff = O.tet4_fff(el, pts)
    n = pts.shape[1]
    x = np.random.default_rng(0).uniform(-1, 1, n)
    yo = np.zeros(n)
    O.lib.oracle_tet4_laplacian_apply(el.shape[1], el, fff, x, yo)
    m = S.Mesh(S.TET4, 1, el, n)
    op = S.laplacian(m, fff)
    assert relerr(host(gpu_apply(op, dev(x))), yo) < RTOL
    do = np.zeros(n)
    O.lib.oracle_tet4_laplacian_diag(el.shape[1], el, fff, do)
    d = zeros(n)
    op.diag(d)
    assert relerr(host(d), do) < RTOL


def test_hex8_laplacian_parity_with_coefficient():
    el, pts = skewed_box(7, perturb=0.03)
    fff = O.hex8_fff(el, pts)
    coeff = np.random.default_rng(1).uniform(0.1, 10, el.shape[1]).astype(np.float32)
    n = pts.shape[1]
    x = np.random.default_rng(0).uniform(-1, 1, n)
    yo = np.zeros(n)
    O.lib.oracle_hex8_laplacian_apply(el.shape[1], el, fff, coeff, x, yo)
    op = S.laplacian(S.Mesh(S.HEX8, 1, el, n), fff, coeff)
    assert relerr(host(gpu_apply(op, dev(x))), yo) < RTOL
    do = np.zeros(n)
    O.lib.oracle_hex8_laplacian_diag(el.shape[1], el, fff, coeff, do)
    d = zeros(n)
    op.diag(d)
    assert relerr(host(d), do) < RTOL


@pytest.mark.parametrize("L,n", [(1, 3), (2, 3), (3, 2), (4, 3), (5, 2), (6, 2), (8, 3), (10, 1), (12, 1), (16, 2)])
def test_sshex8_laplacian_parity(L, n):
    mel, mpts = skewed_box(n, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    fff = O.hex8_fff(mel, mpts)
    coeff = np.random.default_rng(L).uniform(0.1, 10, mel.shape[1]).astype(np.float32)
    N = pts.shape[1]
    x = np.random.default_rng(0).uniform(-1, 1, N)
    m = S.Mesh(S.SSHEX8, L, el, N)
    np.testing.assert_array_equal(m.download(), el)
    for c in (None, coeff):
        yo = np.zeros(N)
        O.lib.oracle_sshex8_laplacian_apply(L, mel.shape[1], el, fff, c, x, yo)
        op = S.laplacian(m, fff, c)
        y0 = dev(np.full(N, 0.5))  # accumulate semantics: y += A x
        op.apply(dev(x), y0)
        assert relerr(host(y0) - 0.5, yo) < RTOL
        do = np.zeros(N)
        O.lib.oracle_sshex8_laplacian_diag(L, mel.shape[1], el, fff, c, do)
        d = zeros(N)
        op.diag(d)
        assert relerr(host(d), do) < RTOL


def test_sshex8_laplacian_fp32_bf16_error_stated():
    L = 8
    mel, mpts = skewed_box(3, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    fff = O.hex8_fff(mel, mpts)
    coeff = np.random.default_rng(2).uniform(0.1, 10, mel.shape[1]).astype(np.float32)
    N = pts.shape[1]
    x = np.random.default_rng(0).uniform(-1, 1, N)
    yo = np.zeros(N)
    O.lib.oracle_sshex8_laplacian_apply(L, mel.shape[1], el, fff, coeff, x, yo)
    op = S.laplacian(S.Mesh(S.SSHEX8, L, el, N), fff, coeff, real_type=S.FLOAT32)
    err = relerr(host(gpu_apply(op, dev(x, np.float32))).astype(np.float64), yo)
    print(f"FP32 apply with BF16 geometry: relative error vs FP64 oracle = {err:.3e}")
    assert err < 1e-2


@pytest.mark.parametrize("L,n", [(1, 3), (2, 3), (3, 2), (4, 3), (5, 2), (6, 2), (8, 3), (10, 1), (12, 1), (16, 2)])
def test_sshex8_laplacian_fp32_every_level(L, n):
    """The FP32 apply (LDS-staged kernel up to L=12, line-thread kernel at L=16) against the FP64
    oracle on the same geometry words (the FP32 operator folds the per-element coefficient into
    the FFF and rounds the products to bf16), accumulate semantics: only the FP32 arithmetic
    differs (< 1e-5); two applies give the same words."""
    mel, mpts = skewed_box(n, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    fff = O.hex8_fff(mel, mpts)
    coeff = np.random.default_rng(L).uniform(0.1, 10, mel.shape[1]).astype(np.float32)
    N = pts.shape[1]
    x = np.random.default_rng(1).uniform(-1, 1, N).astype(np.float32)
    yo = np.zeros(N)
    words = bf16_round((fff.astype(np.float64) * coeff[:, None]).astype(np.float32))
    O.lib.oracle_sshex8_laplacian_apply(L, mel.shape[1], el, words, None, x.astype(np.float64), yo)
    op = S.laplacian(S.Mesh(S.SSHEX8, L, el, N), fff, coeff, real_type=S.FLOAT32)
    outs = []
    for _ in range(2):
        y0 = dev(np.full(N, 0.5, np.float32), np.float32)
        op.apply(dev(x, np.float32), y0)
        S.synchronize()
        outs.append(host(y0))
    err = relerr(outs[0].astype(np.float64) - 0.5, yo)
    print(f"L={L}: FP32 apply vs FP64 oracle on the same bf16 words: {err:.3e}")
    assert err < 1e-5
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("L,n", [(2, 3), (4, 2), (8, 3)])
def test_sshex8_laplacian_atomic_skeleton(L, n):
    """Option "atomic_skeleton": accumulating applies add the skeleton contributions with atomics and
    still match the oracle (FP64) and the two-phase sum (FP32, up to summation order)."""
    mel, mpts = skewed_box(n, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    fff = O.hex8_fff(mel, mpts)
    N = pts.shape[1]
    x = np.random.default_rng(3).uniform(-1, 1, N)
    yo = np.zeros(N)
    O.lib.oracle_sshex8_laplacian_apply(L, mel.shape[1], el, fff, None, x, yo)
    m = S.Mesh(S.SSHEX8, L, el, N)
    op = S.laplacian(m, fff)
    op32 = S.laplacian(m, fff, real_type=S.FLOAT32)
    xd, x32 = dev(x), dev(x, np.float32)
    ref32 = host(gpu_apply(op32, x32)).astype(np.float64)
    y64 = dev(np.full(N, 0.5))  # accumulate semantics: y += A x
    y32 = zeros(N, np.float32)
    S.set_option("atomic_skeleton", 1)
    try:
        op.apply(xd, y64)
        op32.apply(x32, y32)
        S.synchronize()
    finally:
        S.set_option("atomic_skeleton", 0)
    assert relerr(host(y64) - 0.5, yo) < RTOL
    assert relerr(host(y32).astype(np.float64), ref32) < 1e-5


def test_empty_mesh_apply_is_noop():
    el = np.zeros((27, 0), np.int32)
    m = S.Mesh(S.SSHEX8, 2, el, 5)
    op = S.laplacian(m, np.zeros((0, 6), np.float32))
    y = gpu_apply(op, dev(np.ones(5)))
    assert np.all(host(y) == 0)


# ---------------------------------------------------------------- elasticity
# one lane per Gauss point (small meshes), element patches (>= 2^15 elements), one lane per element
@pytest.mark.parametrize("n,patches", [(6, 1), (65, 1), (65, 0)])
def test_hex8_elasticity_parity(n, patches):
    el, pts = skewed_box(n, perturb=0.02 * 6 / n)
    adj, det = O.hex8_adj_det(el, pts)
    coeff = np.random.default_rng(4).uniform(0.1, 10, el.shape[1]).astype(np.float32)
    mu, lam = O.lame(1.0, 0.3)
    n = pts.shape[1] * 3
    x = np.random.default_rng(0).uniform(-1, 1, n)
    S.set_option("hex8_patches", patches)
    try:
        op = S.linear_elasticity(S.Mesh(S.HEX8, 1, el, pts.shape[1]), adj, det, mu, lam, coeff)
        yo = np.full(n, 0.25)
        y = dev(yo)  # accumulate semantics: y += A x
        O.lib.oracle_hex8_elasticity_apply(el.shape[1], el, adj, det, mu, lam, coeff, x, yo)
        op.apply(dev(x), y)
        S.synchronize()
    finally:
        S.set_option("hex8_patches", 1)
    assert relerr(host(y), yo) < RTOL
    do = np.zeros(n)
    O.lib.oracle_hex8_elasticity_diag(el.shape[1], el, adj, det, mu, lam, coeff, do)
    d = zeros(n)
    op.diag(d)
    assert relerr(host(d), do) < RTOL


@pytest.mark.parametrize("n,patches", [(6, 1), (40, 1), (40, 0)])
def test_hex8_elasticity_fp32_bf16_error_stated(n, patches):
    """FP32 hex8 elasticity (bf16 adjugate and determinant words, FP32 sum-factorised arithmetic):
    against the FP64 oracle on the same bf16-rounded words the error is FP32 rounding (< 1e-5); against
    the FP64 oracle on the float32 words it adds the geometry rounding (< 1e-2).  Diagonal and block
    diagonal likewise."""
    el, pts = skewed_box(n, perturb=0.02 * 6 / n)
    adj, det = O.hex8_adj_det(el, pts)
    E = el.shape[1]
    coeff = np.random.default_rng(4).uniform(0.1, 10, E).astype(np.float32)
    mu, lam = O.lame(1.0, 0.3)
    N = pts.shape[1] * 3
    x = np.random.default_rng(0).uniform(-1, 1, N)
    op = S.linear_elasticity(S.Mesh(S.HEX8, 1, el, pts.shape[1]), adj, det, mu, lam, coeff, real_type=S.FLOAT32)
    S.set_option("hex8_patches", patches)
    try:
        y = host(gpu_apply(op, dev(x, np.float32))).astype(np.float64)
    finally:
        S.set_option("hex8_patches", 1)
    yb, yo = np.zeros(N), np.zeros(N)
    O.lib.oracle_hex8_elasticity_apply(E, el, bf16_round(adj), bf16_round(det), mu, lam, coeff, x, yb)
    O.lib.oracle_hex8_elasticity_apply(E, el, adj, det, mu, lam, coeff, x, yo)
    e_arith, e_geom = relerr(y, yb), relerr(y, yo)
    print(f"hex8 elasticity FP32/bf16 n={n}: {e_arith:.2e} vs bf16 words, {e_geom:.2e} vs the FP64 oracle")
    assert e_arith < 1e-5 and e_geom < 1e-2
    db = np.zeros(N)
    O.lib.oracle_hex8_elasticity_diag(E, el, bf16_round(adj), bf16_round(det), mu, lam, coeff, db)
    d = zeros(N, np.float32)
    op.diag(d)
    assert relerr(host(d).astype(np.float64), db) < 1e-6
    b6 = zeros(N // 3 * 6, np.float32)
    op.block_diag_sym(b6)
    b64 = zeros(N // 3 * 6)
    S.linear_elasticity(S.Mesh(S.HEX8, 1, el, pts.shape[1]), bf16_round(adj), bf16_round(det), mu, lam,
                        coeff).block_diag_sym(b64)
    assert relerr(host(b6).astype(np.float64), host(b64)) < 1e-6


@pytest.mark.parametrize("L,n", [(1, 2), (2, 2), (4, 2), (6, 1), (8, 2), (12, 1), (16, 1), (16, 2)])
def test_sshex8_elasticity_parity(L, n):
    mel, mpts = skewed_box(n, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    adj, det = O.hex8_adj_det(mel, mpts)
    coeff = np.exp(np.random.default_rng(L).uniform(np.log(0.1), np.log(10), mel.shape[1])).astype(np.float32)
    mu, lam = O.lame(1.0, 0.3)
    N = pts.shape[1]
    x = np.random.default_rng(0).uniform(-1, 1, 3 * N)
    op = S.linear_elasticity(S.Mesh(S.SSHEX8, L, el, N), adj, det, mu, lam, coeff)
    yo = np.zeros(3 * N)
    O.lib.oracle_sshex8_elasticity_apply(L, mel.shape[1], el, adj, det, mu, lam, coeff, x, yo)
    assert relerr(host(gpu_apply(op, dev(x))), yo) < RTOL
    do = np.zeros(3 * N)
    O.lib.oracle_sshex8_elasticity_diag(L, mel.shape[1], el, adj, det, mu, lam, coeff, do)
    d = zeros(3 * N)
    op.diag(d)
    assert relerr(host(d), do) < RTOL
    bo = np.zeros(6 * N)
    O.lib.oracle_sshex8_elasticity_block_diag(L, mel.shape[1], el, adj, det, mu, lam, coeff, bo)
    b = zeros(6 * N)
    op.block_diag_sym(b)
    S.synchronize()
    assert relerr(host(b), bo) < RTOL
    if L in (8, 16):
        op32 = S.linear_elasticity(S.Mesh(S.SSHEX8, L, el, N), adj, det, mu, lam, coeff, real_type=S.FLOAT32)
        err = relerr(host(gpu_apply(op32, dev(x, np.float32))).astype(np.float64), yo)
        print(f"FP32 elasticity apply with BF16 geometry: relative error vs FP64 oracle = {err:.3e}")
        assert err < 1e-2


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("L,n", [(2, 2), (3, 2), (4, 2), (5, 1), (6, 2), (8, 2), (10, 1), (12, 1), (16, 2)])
def test_sshex8_elasticity_stencil_and_element_forms(L, n, form):
    """Both forms of the sshex8 elasticity apply (1: element form on the matrix cores / node-wise
    kernel, 2: the 27-point 3x3-block stencil form) against the FP64 oracle, accumulating into a
    nonzero y, with per-element coefficients; FP32 stencil form on the same BF16 geometry words."""
    mel, mpts = skewed_box(n, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    adj, det = O.hex8_adj_det(mel, mpts)
    E = mel.shape[1]
    coeff = np.exp(np.random.default_rng(L + 7).uniform(np.log(0.1), np.log(10), E)).astype(np.float32)
    mu, lam = O.lame(1.0, 0.3)
    N = pts.shape[1]
    rng = np.random.default_rng(L)
    x = rng.uniform(-1, 1, 3 * N)
    y0 = rng.uniform(-1, 1, 3 * N)
    S.set_option("elasticity_form", form)
    try:
        mesh = S.Mesh(S.SSHEX8, L, el, N)
        op = S.linear_elasticity(mesh, adj, det, mu, lam, coeff)
        yo = y0.copy()
        O.lib.oracle_sshex8_elasticity_apply(L, E, el, adj, det, mu, lam, coeff, x, yo)
        y = dev(y0)
        op.apply(dev(x), y)
        S.synchronize()
        err = relerr(host(y), yo)
        print(f"L={L} form={form}: FP64 accumulate vs oracle {err:.2e}")
        assert err < RTOL
        if form == 2 and sshex8_stencil(L):
            x32 = x.astype(np.float32)
            op32 = S.linear_elasticity(mesh, adj, det, mu, lam, coeff, real_type=S.FLOAT32)
            yb = np.zeros(3 * N)
            O.lib.oracle_sshex8_elasticity_apply(L, E, el, bf16_round(adj), bf16_round(det), mu, lam, coeff,
                                                 x32.astype(np.float64), yb)
            y1 = host(gpu_apply(op32, dev(x32, np.float32)))
            y2 = host(gpu_apply(op32, dev(x32, np.float32)))
            assert np.array_equal(y1, y2)
            err32 = relerr(y1.astype(np.float64), yb)
            print(f"L={L}: FP32 stencil form vs FP64 oracle on the same BF16 geometry {err32:.2e}")
            assert err32 < 1e-5
    finally:
        S.set_option("elasticity_form", 0)


def sshex8_stencil(L):
    return L in (2, 3, 4, 5, 6, 8, 10, 12, 16)


def bf16_round(a):
    """float32 -> nearest-even bfloat16, returned as float32 (the FP32 operator's geometry words)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("L,n", [(2, 3), (4, 3), (4, 2), (6, 2), (8, 2), (12, 1), (16, 1)])
def test_sshex8_elasticity_fp32_matrix_cores(L, n):
    """The FP32 sshex8 elasticity operator runs on the matrix cores (bf16 hi/lo split products,
    FP32 accumulation).  On the same BF16-rounded geometry as the FP64 oracle it must agree to the
    splitting error (~1e-5), and repeated applies must be bitwise identical."""
    mel, mpts = skewed_box(n, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    adj, det = O.hex8_adj_det(mel, mpts)
    E = mel.shape[1]
    coeff = np.exp(np.random.default_rng(L + 1).uniform(np.log(0.1), np.log(10), E)).astype(np.float32)
    mu, lam = O.lame(1.0, 0.3)
    N = pts.shape[1]
    x = np.random.default_rng(3).uniform(-1, 1, 3 * N).astype(np.float32)
    op32 = S.linear_elasticity(S.Mesh(S.SSHEX8, L, el, N), adj, det, mu, lam, coeff, real_type=S.FLOAT32)
    yo = np.zeros(3 * N)
    O.lib.oracle_sshex8_elasticity_apply(L, E, el, bf16_round(adj), bf16_round(det), mu, lam, coeff,
                                         x.astype(np.float64), yo)
    y1 = host(gpu_apply(op32, dev(x, np.float32)))
    y2 = host(gpu_apply(op32, dev(x, np.float32)))
    assert np.array_equal(y1, y2)
    err = relerr(y1.astype(np.float64), yo)
    print(f"L={L}: FP32 matrix-core elasticity apply vs FP64 oracle (same geometry) = {err:.3e}")
    assert err < 1e-4


# ---------------------------------------------------------------- stored element matrices (ss:em)
@pytest.mark.parametrize("L,n,bs", [(2, 2, 1), (8, 2, 1), (16, 1, 1), (4, 2, 3), (16, 2, 3)])
def test_sshex8_element_matrix_parity(L, n, bs):
    mel, mpts = skewed_box(n, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    E = mel.shape[1]
    rng = np.random.default_rng(L * 10 + bs)
    K = rng.normal(size=(E, (8 * bs) ** 2))  # non-symmetric: pins the row/column convention
    N = pts.shape[1] * bs
    x = rng.uniform(-1, 1, N)
    yo = np.zeros(N)
    O.lib.oracle_sshex8_em_apply(L, E, el, bs, K, x, yo)
    mesh = S.Mesh(S.SSHEX8, L, el, pts.shape[1])
    op = S.element_matrix_op(mesh, K, bs)
    assert relerr(host(gpu_apply(op, dev(x))), yo) < RTOL
    op32 = S.element_matrix_op(mesh, K, bs, real_type=S.FLOAT32)
    err = relerr(host(gpu_apply(op32, dev(x, np.float32))).astype(np.float64), yo)
    print(f"L={L} bs={bs}: FP32 ss:em apply relative error vs FP64 oracle = {err:.3e}")
    assert err < 1e-5


# ---------------------------------------------------------------- transfers
@pytest.mark.parametrize("vs", [1, 3])
def test_transfers_parity(vs):
    L = 8
    mel, mpts = skewed_box(3, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    E = mel.shape[1]
    fine = S.Mesh(S.SSHEX8, L, el, pts.shape[1])
    count = O.incidence(el, pts.shape[1])
    rng = np.random.default_rng(7)
    prev_el, prev_mesh, prev_L = el, fine, L
    for Lc in (4, 2, 1):
        coarse = prev_mesh.derefine(Lc)
        ec = O.derefine(L, Lc, el)
        if Lc == 1:
            np.testing.assert_array_equal(coarse.download(), O.to_hex8(1, ec))
        else:
            np.testing.assert_array_equal(coarse.download(), ec)
        nc, nf = coarse.nnodes, prev_mesh.nnodes
        assert nc == ec.max() + 1
        xc = rng.uniform(-1, 1, nc * vs)
        yo = np.zeros(nf * vs)
        if Lc == 1:
            O.lib.oracle_sshex8_hierarchical_prolongation(prev_L, E, prev_el, vs, xc, yo)
        else:
            O.lib.oracle_sshex8_prolongate(E, Lc, ec, prev_L, prev_el, vs, xc, yo)
        P = S.prolongation(coarse, prev_mesh, vs)
        y = zeros(nf * vs)
        P.apply(dev(xc), y)
        S.synchronize()
        assert relerr(host(y), yo) < RTOL
        xf = rng.uniform(-1, 1, nf * vs)
        cnt = O.incidence(prev_el, nf)
        ro = np.zeros(nc * vs)
        if Lc == 1:
            O.lib.oracle_sshex8_hierarchical_restriction(prev_L, E, prev_el, cnt, vs, xf, ro)
        else:
            O.lib.oracle_sshex8_restrict(E, prev_L, prev_el, cnt, Lc, ec, vs, xf, ro)
        R = S.restriction(prev_mesh, coarse, vs)
        r = zeros(nc * vs)
        R.apply(dev(xf), r)
        S.synchronize()
        assert relerr(host(r), ro) < RTOL
        prev_el, prev_mesh, prev_L = ec, coarse, Lc
    del count


@pytest.mark.parametrize("vs", [1, 3])
def test_transfers_level16_parity(vs):
    """Level 16 -> 8: the restriction's element tile exceeds the LDS budget, so it runs on the 2^3
    sub-blocks of level 8 (onto level 4); the prolongation stages the coarse level-8 tile."""
    L, Lc = 16, 8
    mel, mpts = skewed_box(2, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    E, nf = mel.shape[1], pts.shape[1]
    fine = S.Mesh(S.SSHEX8, L, el, nf)
    coarse = fine.derefine(Lc)
    ec = O.derefine(L, Lc, el)
    np.testing.assert_array_equal(coarse.download(), ec)
    nc = coarse.nnodes
    rng = np.random.default_rng(11)
    xc = rng.uniform(-1, 1, nc * vs)
    yo = np.zeros(nf * vs)
    O.lib.oracle_sshex8_prolongate(E, Lc, ec, L, el, vs, xc, yo)
    xf = rng.uniform(-1, 1, nf * vs)
    ro = np.zeros(nc * vs)
    O.lib.oracle_sshex8_restrict(E, L, el, O.incidence(el, nf), Lc, ec, vs, xf, ro)
    for rt, dt, tol in ((S.FLOAT64, np.float64, RTOL), (S.FLOAT32, np.float32, 1e-6)):
        P = S.prolongation(coarse, fine, vs, real_type=rt)
        y = zeros(nf * vs, dt)
        P.apply(dev(xc, dt), y)
        R = S.restriction(fine, coarse, vs, real_type=rt)
        r = zeros(nc * vs, dt)
        R.apply(dev(xf, dt), r)
        S.synchronize()
        assert relerr(host(y).astype(np.float64), yo) < tol
        assert relerr(host(r).astype(np.float64), ro) < tol


def test_constrained_operator_copies_dofs():
    L = 4
    mel, mpts = skewed_box(2)
    el, pts = O.sshex8_refine(L, mel, mpts)
    fff = O.hex8_fff(mel, mpts)
    N = pts.shape[1]
    nodes = np.where(pts[0] < 1e-6)[0].astype(np.int32)
    op = S.laplacian(S.Mesh(S.SSHEX8, L, el, N), fff)
    cop = S.constrained(op, nodes, 1)
    x = np.random.default_rng(0).uniform(-1, 1, N)
    yo = np.zeros(N)
    O.lib.oracle_sshex8_laplacian_apply(L, mel.shape[1], el, fff, None, x, yo)
    O.lib.oracle_constraint_nodes_copy_vec(len(nodes), nodes, 1, 0, x, yo)
    assert relerr(host(gpu_apply(cop, dev(x))), yo) < RTOL
    z = dev(np.ones(N))
    S.set_constrained_value(cop, 0.0, z)
    S.synchronize()
    zz = host(z)
    assert np.all(zz[nodes] == 0) and zz.sum() == N - len(nodes)


@pytest.mark.parametrize("kind,L", [("laplacian", 4), ("elasticity", 4), ("elasticity", 16)])
@pytest.mark.parametrize("interior", [False, True])
def test_constrained_gather_matches_separate_copy(kind, L, interior):
    """Dirichlet rows on the macro-element skeleton are written by the operator's skeleton gather
    (y = x there); with a constrained interior dof the operator falls back to apply + copy.  Both
    must equal the unconstrained apply followed by the copy, bitwise (accumulating into y)."""
    mel, mpts = skewed_box(2, perturb=0.02)
    el, pts = O.sshex8_refine(L, mel, mpts)
    N = pts.shape[1]
    mesh = S.Mesh(S.SSHEX8, L, el, N)
    nodes = np.where(pts[0] < 1e-6)[0].astype(np.int32)
    if interior:
        nodes = np.append(nodes, el[(L // 2 * (L + 1) + L // 2) * (L + 1) + L // 2, 0]).astype(np.int32)
    rng = np.random.default_rng(L)
    for rt, dt in ((S.FLOAT64, np.float64), (S.FLOAT32, np.float32)):
        if kind == "laplacian":
            bs, comps = 1, None
            op = S.laplacian(mesh, O.hex8_fff(mel, mpts), real_type=rt)
        else:
            bs, comps = 3, [0, 2]
            adj, det = O.hex8_adj_det(mel, mpts)
            mu, lam = O.lame(1.0, 0.3)
            op = S.linear_elasticity(mesh, adj, det, mu, lam, real_type=rt)
        cop = S.constrained(op, nodes, bs, comps)
        x = rng.uniform(-1, 1, bs * N).astype(dt)
        y0 = rng.uniform(-1, 1, bs * N).astype(dt)
        yr = dev(y0, dt)
        op.apply(dev(x, dt), yr)
        ref = host(yr)
        dofs = (nodes[:, None] * bs + np.asarray(comps if comps else [0])[None, :]).ravel()
        ref[dofs] = x[dofs]
        y = dev(y0, dt)
        cop.apply(dev(x, dt), y)
        S.synchronize()
        np.testing.assert_array_equal(host(y), ref)


# ---------------------------------------------------------------- BLAS
@pytest.mark.parametrize("rt", [S.FLOAT64, S.FLOAT32])
def test_fused_vector_steps(rt):
    """sfem_amd_cheb3_start (Chebyshev3's first step, with and without A x and D^-1) and
    sfem_amd_cg_update_x_r_dot (CG's x/r update returning (r,r) and (r,z)) against numpy."""
    dt = np.float64 if rt == S.FLOAT64 else np.float32
    tol = 1e-13 if rt == S.FLOAT64 else 1e-6
    rng = np.random.default_rng(4)
    n = 100_003
    rhs, ax, dinv, x0 = (rng.uniform(-1, 1, n).astype(dt) for _ in range(4))
    alpha = 0.37
    for have_ax in (0, 1):
        for use_d in (False, True):
            p = dev(ax, dt)
            x = dev(x0, dt)
            d = dev(dinv, dt) if use_d else None
            b = dev(rhs, dt)  # device arrays stay referenced until the kernel has run
            assert S.lib.sfem_amd_cheb3_start(rt, n, alpha, S._dptr(b), S._dptr(d), have_ax, S._dptr(p), S._dptr(x),
                                              None) == 0
            S.synchronize()
            po = (ax.astype(np.float64) - rhs) if have_ax else -rhs.astype(np.float64)
            if use_d:
                po = dinv * po
            xo = (x0 if have_ax else 0) - alpha * po
            assert relerr(host(p).astype(np.float64), po) < tol
            assert relerr(host(x).astype(np.float64), xo) < tol
    pv, apv, xv, rv, zv = (rng.uniform(-1, 1, n).astype(dt) for _ in range(5))
    x, r, p, ap, z = dev(xv, dt), dev(rv, dt), dev(pv, dt), dev(apv, dt), dev(zv, dt)
    rz = C.c_double()
    rr = S.lib.sfem_amd_cg_update_x_r_dot(rt, n, alpha, S._dptr(p), S._dptr(ap), S._dptr(x), S._dptr(r), S._dptr(z),
                                          C.byref(rz), None)
    xo = xv.astype(np.float64) + alpha * pv
    ro = rv.astype(np.float64) - alpha * apv
    assert relerr(host(x).astype(np.float64), xo) < tol and relerr(host(r).astype(np.float64), ro) < tol
    assert abs(rr - ro @ ro) < tol * (ro @ ro)
    assert abs(rz.value - ro @ zv) < tol * (np.abs(ro) @ np.abs(zv))


def test_blas_reductions_deterministic_and_accurate():
    rng = np.random.default_rng(0)
    for n in (0, 1, 1000, 3_000_001):
        a, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        da, db = dev(a), dev(b)
        d1 = S.dot(da, db)
        d2 = S.dot(da, db)
        assert d1 == d2  # bitwise reproducible
        assert abs(d1 - a @ b) <= 1e-12 * max(1.0, np.abs(a) @ np.abs(b))
        if n:
            assert abs(S.norm2(da) - np.linalg.norm(a)) < 1e-12 * np.linalg.norm(a)


@pytest.mark.parametrize("offset", [0, 1, 4])
def test_fp32_streaming_kernels_aligned_and_not(offset):
    """FP32 axpby, fill, copy and the Chebyshev3 steps take 16 bytes per thread and array when every
    array is 16-byte aligned (the last n % 4 elements by the first threads of block 0) and one
    float per thread otherwise: both against numpy in FP32 arithmetic, on arrays starting `offset`
    floats into their allocation (4: aligned, 1: not), with n = 1001."""
    n = 1001
    rng = np.random.default_rng(offset)
    host_in = {k: rng.uniform(-1, 1, n + 8).astype(np.float32) for k in ("x", "y", "rhs", "ax", "dinv", "p", "z")}
    arr = {k: dev(v, np.float32) for k, v in host_in.items()}
    ptr = {k: C.c_void_p(a.ptr + 4 * offset) for k, a in arr.items()}
    sl = slice(offset, offset + n)
    f = np.float32
    a, b = f(0.75), f(-1.25)
    assert S.lib.sfem_amd_blas_axpby(S.FLOAT32, n, float(a), ptr["x"], float(b), ptr["y"], None) == 0
    assert S.lib.sfem_amd_blas_copy(S.FLOAT32, n, ptr["x"], ptr["z"], None) == 0
    assert S.lib.sfem_amd_blas_values(S.FLOAT32, n, 2.5, ptr["dinv"], None) == 0
    alpha, beta = f(0.3), f(0.6)
    assert S.lib.sfem_amd_cheb3_update(S.FLOAT32, n, float(alpha), float(beta), ptr["rhs"], ptr["ax"], ptr["p"],
                                       ptr["x"], None) == 0
    S.synchronize()
    got = {k: host(v) for k, v in arr.items()}
    x0, y0, rhs, ax, p0 = (host_in[k][sl] for k in ("x", "y", "rhs", "ax", "p"))
    np.testing.assert_allclose(got["y"][sl], a * x0 + b * y0, rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(got["z"][sl], x0)
    np.testing.assert_array_equal(got["dinv"][sl], np.full(n, 2.5, np.float32))
    pn = (beta * p0 + (ax - rhs)).astype(np.float32)  # fma(beta, p, Ax - rhs) in FP32
    np.testing.assert_allclose(got["p"][sl], pn, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(got["x"][sl], x0 - alpha * got["p"][sl], rtol=1e-6, atol=1e-7)
    for k in ("x", "y", "z", "dinv", "p"):  # nothing outside [offset, offset + n) written
        np.testing.assert_array_equal(got[k][:offset], host_in[k][:offset])
        np.testing.assert_array_equal(got[k][offset + n:], host_in[k][offset + n:])
