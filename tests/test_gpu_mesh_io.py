"""A mesh read from a raw mesh folder written by the reference's own box-mesh writer
(tests/golden/raw_hex8_2x2x2: meta.yaml, x/y/z.raw, i{k}.raw, boundary node sets) driven through the
HIP path: the hex8 mesh and its sshex8 refinement applied on the GPU against the oracle, and a
Dirichlet problem on a boundary node set read from the folder solved with MG-PCG against the
oracle's CG solve."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

import sfem_amd as S  # noqa: E402
from sfem_amd import gmg, mesh_io  # noqa: E402

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import solvers as SO  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "raw_hex8_2x2x2")


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _apply(op, x, dtype=np.float64):
    y = S.DeviceArray(op.rows, dtype)
    op.apply(S.DeviceArray.from_numpy(x, dtype), y)
    S.synchronize()
    return y.numpy()


@pytest.mark.parametrize("L", [1, 4])
def test_read_folder_apply_matches_oracle(L):
    m = mesh_io.read_mesh(GOLDEN)
    mel, mpts = m["elements"], m["points"]
    E = mel.shape[1]
    fff = O.hex8_fff(mel, mpts)
    adj, det = O.hex8_adj_det(mel, mpts)
    mu, lam = O.lame(1.0, 0.3)
    if L == 1:
        mesh, N = S.Mesh(S.HEX8, 1, mel, mpts.shape[1]), mpts.shape[1]
    else:
        el, pts = O.sshex8_refine(L, mel, mpts)
        mesh, N = S.Mesh(S.SSHEX8, L, el, pts.shape[1]), pts.shape[1]
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, N)
    yo = np.zeros(N)
    if L == 1:
        O.lib.oracle_hex8_laplacian_apply(E, mel, fff, None, x, yo)
    else:
        O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, None, x, yo)
    assert _rel(_apply(S.laplacian(mesh, fff), x), yo) < 1e-12
    x3 = rng.uniform(-1, 1, 3 * N)
    yo3 = np.zeros(3 * N)
    if L == 1:
        O.lib.oracle_hex8_elasticity_apply(E, mel, adj, det, mu, lam, None, x3, yo3)
    else:
        O.lib.oracle_sshex8_elasticity_apply(L, E, el, adj, det, mu, lam, None, x3, yo3)
    assert _rel(_apply(S.linear_elasticity(mesh, adj, det, mu, lam), x3), yo3) < 1e-12


def test_read_folder_dirichlet_solve_matches_oracle():
    """Poisson, level 4 on the folder's mesh, zero Dirichlet on its "bottom" and "top" node sets (refined
    nodes on those faces included), f = 1: MG-PCG on the GPU against the oracle's Jacobi-PCG solve."""
    L = 4
    m = mesh_io.read_mesh(GOLDEN)
    mel, mpts = m["elements"], m["points"]
    E = mel.shape[1]
    el, pts = O.sshex8_refine(L, mel, mpts)
    N = pts.shape[1]
    fff = O.hex8_fff(mel, mpts)
    # the folder's node sets name macro nodes; the faces they span give the refined ones
    faces = []
    for name in ("bottom", "top"):
        macro = m["boundary_nodes"][name]
        for k in range(3):
            v = mpts[k, macro]
            if np.all(v == v[0]):
                faces.append((k, v[0]))
    assert len(faces) == 2
    nodes = np.where(np.any([np.abs(pts[k] - v) <= 1e-6 * max(1.0, abs(float(v))) for k, v in faces],
                            axis=0))[0].astype(np.int32)
    b = np.ones(N)
    b[nodes] = 0
    h = gmg.SSGMG(S.Mesh(S.SSHEX8, L, el, N), "laplacian", fff, nodes, coarse_rtol=1e-12)
    solver = S.cg(h.fine.A, h.mg, rtol=1e-12, max_it=200)
    x = S.DeviceArray(N)
    solver.apply(S.DeviceArray.from_numpy(b), x)

    def A_o(u, y):
        O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, None, u, y)
        y[nodes] = u[nodes]
    d = np.zeros(N)
    O.lib.oracle_sshex8_laplacian_diag(L, E, el, fff, None, d)
    d[nodes] = 1
    xo, it_o, st_o, _ = SO.cg(A_o, b, np.zeros(N), rtol=1e-13, max_it=5000, precond=lambda r, z: z.__iadd__(r / d))
    diff = _rel(x.numpy(), xo)
    print(f"raw-folder mesh, L={L}: MG-PCG {solver.iterations} its, difference to the oracle solve {diff:.2e}")
    assert solver.last_status == 0 and st_o == 0
    assert diff < 1e-9
