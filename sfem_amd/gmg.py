"""Geometric multigrid on sshex8 meshes, built from the library's device objects.

Mirrors create_ssgmg / create_gmg_data (ssmg/sfem_ssgmg.cpp:9-42,
algebra/mg/sfem_GeometricMultigrid.cpp:12-168) with the smoothing of drivers/mgsolve.cpp:
  * levels = derefinement levels of L (L, L/2, ... down to the odd part, then 1; 8 -> 8,4,2,1);
    level 1 is the HEX8 macro mesh;
  * every level rediscretises the operator with the macro geometry (dirichlet_as_zero on
    coarse levels: the coarse constrained nodes are the fine ones that exist there);
  * prolongation and restriction zero the constrained dofs of their output level;
  * fine levels: Chebyshev3 (degree 3, power method from ones with constrained dofs zeroed,
    scale_eig_max 1.02, tolerance 1e-5, drivers/mgsolve.cpp:68-71,176-185), or the
    reference's default damped Jacobi (create_gmg_default_smoothers_and_solver);
  * coarsest: FP64 CG with inverse-diagonal preconditioning (drivers/mgsolve.cpp:214-227).
"""
import numpy as np

from . import FLOAT32, FLOAT64, HEX8, SSHEX8

from . import (DeviceArray, Mesh, Multigrid, cg, cheb3, constrained, inverse_diagonal, jacobi, laplacian,
               linear_elasticity, precision_bridge, prolongation, restriction, set_constrained_value)


def derefinement_levels(L):
    """[L, L/2, ..., m, 1] with m the odd part of L (frontend/tests/sfem_SemiStructuredMeshTest.cpp:18-22)."""
    out = [L]
    while out[-1] % 2 == 0 and out[-1] > 1:
        out.append(out[-1] // 2)
    if out[-1] != 1:
        out.append(1)
    return out


class Level:
    pass


class SSGMG:
    """Holds the hierarchy; `mg` is the Multigrid operator (one V-cycle per apply when used
    as a preconditioner).  kind: "laplacian" or "elasticity"."""

    def __init__(self, fine_mesh, kind, geometry, bc_nodes, coeff=None, real_type=FLOAT64, smoother="cheb3",
                 cheb_degree=3, eig_its=10, eig_tol=1e-5, scale_eig_max=1.02, coarse_rtol=1e-8,
                 coarse_max_it=10000, mu=None, lam=None, bc_components=None, max_it=1, atol=0.0):
        self.kind = kind
        self.block_size = 3 if kind == "elasticity" else 1
        self.real_type = real_type
        dtype = np.float64 if real_type == FLOAT64 else np.float32
        bs = self.block_size
        bc_nodes = np.asarray(bc_nodes, np.int32)
        self.levels = []
        meshes = [fine_mesh]
        for l in derefinement_levels(fine_mesh.level)[1:]:
            meshes.append(meshes[-1].derefine(l))
        for i, m in enumerate(meshes):
            lv = Level()
            lv.mesh = m
            if kind == "laplacian":
                lv.op = laplacian(m, geometry, coeff, real_type=real_type if m.element_type == SSHEX8 else FLOAT64)
            else:
                adj, det = geometry
                lv.op = linear_elasticity(m, adj, det, mu, lam, coeff,
                                          real_type=real_type if m.element_type == SSHEX8 else FLOAT64)
            lv.bc = bc_nodes[bc_nodes < m.nnodes]
            lv.A = constrained(lv.op, lv.bc, bs, bc_components)
            lv.n = m.nnodes * bs
            self.levels.append(lv)
        self.coarse_real_type = FLOAT64
        # fine-level smoothers
        for i, lv in enumerate(self.levels[:-1]):
            if smoother == "cheb3":
                init = DeviceArray.from_numpy(np.ones(lv.n, dtype))
                set_constrained_value(lv.A, 0.0, init)
                lv.smoother = cheb3(lv.A, max_it=cheb_degree, scale_eig_max=scale_eig_max, eigen_solver_max_it=eig_its,
                                    eigen_solver_tol=eig_tol, init_vector=init)
            else:  # the reference default: 3 sweeps of damped Jacobi
                d = DeviceArray(lv.n, dtype)
                lv.A.diag(d)
                lv.smoother = jacobi(lv.A, d, omega=1.0 / bs, steps=3)
        # coarsest level: FP64 PCG
        c = self.levels[-1]
        d = DeviceArray(c.n, np.float64)
        c.A.diag(d)
        c.precond = inverse_diagonal(d, 1.0)
        c.smoother = cg(c.A, c.precond, rtol=coarse_rtol, atol=1e-30, max_it=coarse_max_it)
        # transfers (zero constraints on the output level)
        for i in range(len(self.levels) - 1):
            f, co = self.levels[i], self.levels[i + 1]
            rt = real_type
            f.R = restriction(f.mesh, co.mesh, bs, rt, constraints=co.A)
            co.P = prolongation(co.mesh, f.mesh, bs, rt, constraints=f.A)
        self.mg = Multigrid(real_type, 1, max_it, atol)
        for i, lv in enumerate(self.levels):
            A, sm = lv.A, lv.smoother
            if real_type == FLOAT32 and i == len(self.levels) - 1:
                # the coarse solve stays FP64 inside the FP32 cycle
                A, sm = precision_bridge(A), precision_bridge(sm)
                lv.bridges = (A, sm)
            self.mg.add_level(A, sm, getattr(lv, "P", None), getattr(lv, "R", None))

    @property
    def fine(self):
        return self.levels[0]
