"""The C++ drop-in (include/sfem/: create_ssgmg, create_cg, Multigrid, ConjugateGradient over the C ABI)
built with g++ as a user of the reference's C++ API would build it, and its MG-PCG solve compared with
the oracle's restatement of the reference's default geometric multigrid
(algebra/mg/sfem_GeometricMultigrid.cpp:12-168, ssmg/sfem_ssgmg.cpp:9-42)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import solvers as SO  # noqa: E402

LIBDIR = os.path.join(ROOT, "sfem_amd", "lib")


def build_example(out_dir):
    exe = os.path.join(str(out_dir), "drop_in_example")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "drop_in_example.cpp"), "-L", LIBDIR, "-lsfem_amd",
           f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_cpp_drop_in_compiles_against_the_headers(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libsfem_amd.so")):
        pytest.skip("libsfem_amd.so not built")
    assert os.path.exists(build_example(tmp_path))


def _constrained(apply, dofs):
    def f(x, y):
        apply(x, y)
        y[dofs] = x[dofs]
    return f


def oracle_default_gmg(kind, L, mel, mpts, el, nodes, geom, mu=0.0, lam=0.0):
    """The reference's default hierarchy restated on the oracle: levels L, L/2, ..., 1; Jacobi
    (omega 1, scalar) or symmetric 3x3 block Jacobi (omega 1/3, vector) with 3 sweeps; coarse PCG
    with rtol 1e-6 preconditioned by the inverse diagonal / block size."""
    E = mel.shape[1]
    bs = 3 if kind == "elasticity" else 1
    levels = O.hierarchical_levels(L)[::-1]
    conns = [el] + [O.derefine(L, l, el) for l in levels[1:]]
    out = []
    for i, (l, c) in enumerate(zip(levels, conns)):
        nn = int(c.max()) + 1
        n = nn * bs
        bc = nodes[nodes < nn]
        dofs = (bc[:, None].astype(np.int64) * bs + np.arange(bs)).ravel()
        h8 = O.to_hex8(1, c) if l == 1 else None
        if kind == "laplacian":
            fff = geom
            if l == 1:
                ap = (lambda h8: lambda x, y: O.lib.oracle_hex8_laplacian_apply(E, h8, fff, None, x, y))(h8)
            else:
                ap = (lambda l, c: lambda x, y: O.lib.oracle_sshex8_laplacian_apply(l, E, c, fff, None, x, y))(l, c)
        else:
            adj, det = geom
            if l == 1:
                ap = (lambda h8: lambda x, y: O.lib.oracle_hex8_elasticity_apply(E, h8, adj, det, mu, lam, None,
                                                                                 x, y))(h8)
            else:
                ap = (lambda l, c: lambda x, y: O.lib.oracle_sshex8_elasticity_apply(l, E, c, adj, det, mu, lam,
                                                                                     None, x, y))(l, c)
        A = _constrained(ap, dofs)
        if i < len(levels) - 1:
            if bs == 1:
                d = np.zeros(n)
                O.lib.oracle_sshex8_laplacian_diag(l, E, c, fff, None, d)
                d[dofs] = 1
                sm = SO.Jacobi(A, d, omega=1.0, steps=3)
            else:
                sym = np.zeros(6 * nn)
                O.lib.oracle_sshex8_elasticity_block_diag(l, E, c, adj, det, mu, lam, None, sym)
                mask = np.zeros(n, np.uint8)
                mask[dofs] = 1
                sm = SO.BlockJacobi(A, sym, mask, omega=1.0 / 3, steps=3)
        else:
            d = np.zeros(n)
            if bs == 1:
                O.lib.oracle_hex8_laplacian_diag(E, h8, fff, None, d)
            else:
                O.lib.oracle_hex8_elasticity_diag(E, h8, adj, det, mu, lam, None, d)
            d[dofs] = 1

            def sm(b, x, A=A, d=d):
                x[:] = SO.cg(A, b, np.zeros_like(b), rtol=1e-6, atol=1e-16, max_it=10000,
                             precond=lambda r, z: z.__iadd__(r / (bs * d)))[0]
                return x
        out.append([A, sm, None, None, n, dofs, c, l, nn])
    for i in range(len(out) - 1):
        f, co = out[i], out[i + 1]
        cnt = O.incidence(f[6], f[8])

        def R(r, f=f, co=co, cnt=cnt):
            rc = np.zeros(co[4])
            if co[7] == 1:
                O.lib.oracle_sshex8_hierarchical_restriction(f[7], E, f[6], cnt, bs, r, rc)
            else:
                O.lib.oracle_sshex8_restrict(E, f[7], f[6], cnt, co[7], co[6], bs, r, rc)
            rc[co[5]] = 0
            return rc

        def P(xc, f=f, co=co):
            xf = np.zeros(f[4])
            if co[7] == 1:
                O.lib.oracle_sshex8_hierarchical_prolongation(f[7], E, f[6], bs, xc, xf)
            else:
                O.lib.oracle_sshex8_prolongate(E, co[7], co[6], f[7], f[6], bs, xc, xf)
            xf[f[5]] = 0
            return xf
        f[3] = R
        co[2] = P
    return SO.Multigrid([tuple(o[:4]) for o in out], max_it=1, atol=0), out[0][0]


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["laplacian", "elasticity"])
def test_cpp_create_ssgmg_mg_pcg_matches_oracle(kind, tmp_path):
    """drop_in_example (N = 3 macro elements per axis, L = 4): CG preconditioned by one V-cycle of
    create_ssgmg's default hierarchy against the oracle's default hierarchy in the oracle's PCG on
    the same problem; iteration counts within one, solutions to 1e-8, oracle-operator residual."""
    n, L = 3, 4
    exe = build_example(tmp_path)
    out = os.path.join(str(tmp_path), "x.bin")
    print(f"drop_in_example {kind} {n} {L}", flush=True)
    r = subprocess.run([exe, kind, str(n), str(L), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    its = int(r.stdout.split()[1])
    x = np.fromfile(out, np.float64)

    mel, mpts = O.box_mesh(n + 1, n + 1, n + 1)
    el, pts = O.sshex8_refine(L, mel, mpts)
    N = pts.shape[1]
    if kind == "laplacian":
        on = (pts < 1e-6) | (pts > 1 - 1e-6)
        nodes = np.where(on.any(axis=0))[0].astype(np.int32)
        geom, mu, lam = O.hex8_fff(mel, mpts), 0.0, 0.0
        b = np.ones(N)
    else:
        nodes = np.where(pts[2] < 1e-6)[0].astype(np.int32)
        geom = O.hex8_adj_det(mel, mpts)
        mu, lam = O.lame(1.0, 0.3)
        b = np.zeros((N, 3))
        b[:, 2] = -1
        b = b.ravel()
    bs = 3 if kind == "elasticity" else 1
    b[(nodes[:, None].astype(np.int64) * bs + np.arange(bs)).ravel()] = 0
    mg, A = oracle_default_gmg(kind, L, mel, mpts, el, nodes, geom, mu, lam)
    xo, it_o, st_o, _ = SO.cg(A, b, np.zeros_like(b), rtol=1e-12, max_it=500,
                              precond=lambda r, z: z.__iadd__(mg(r, np.zeros_like(r))))
    Ax = np.zeros_like(b)
    A(x, Ax)
    res = np.linalg.norm(b - Ax) / np.linalg.norm(b)
    diff = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"C++ create_ssgmg {kind}: {its} its (oracle {it_o}), residual {res:.2e}, difference {diff:.2e}")
    assert st_o == 0
    assert abs(its - it_o) <= 1
    assert res < 1e-10
    assert diff < 1e-8


def test_cpp_operator_algebra_semantics(tmp_path):
    """tests/cpp/operator_algebra.cpp (host lambdas only): OperatorAdd, OperatorCompose and an
    InPlaceOperator composed after an operator follow the reference's rules."""
    if not os.path.exists(os.path.join(LIBDIR, "libsfem_amd.so")):
        pytest.skip("libsfem_amd.so not built")
    exe = os.path.join(str(tmp_path), "operator_algebra")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "operator_algebra.cpp"), "-L", LIBDIR, "-lsfem_amd",
           f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


def test_cpp_lanczos_closed_forms(tmp_path):
    """tests/cpp/lanczos_host.cpp (host BLAS): the Lanczos bound of include/sfem/sfem_Lanczos.hpp
    against closed-form spectra: the tridiagonal bisection, the 1-D Laplacian (to convergence and
    as a lower bound when cut short), D^-1 A similar to it, and decoupled rows left out by a zero
    start entry."""
    exe = os.path.join(str(tmp_path), "lanczos_host")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "lanczos_host.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
