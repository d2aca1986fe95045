"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) for the tests.

Test infrastructure: the oracle is the parity checker, never the product.
All arrays are numpy, C-contiguous; connectivity is SoA int32 [nxe][E].
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C")
_pd = C.c_ssize_t


class _NullableF32:
    """Accepts None (NULL) or a float32 array."""

    @classmethod
    def from_param(cls, obj):
        if obj is None:
            return None
        a = np.ascontiguousarray(obj, dtype=np.float32)
        return a.ctypes.data_as(C.c_void_p)


class _NullableI32:
    @classmethod
    def from_param(cls, obj):
        if obj is None:
            return None
        a = np.ascontiguousarray(obj, dtype=np.int32)
        return a.ctypes.data_as(C.c_void_p)


def _load():
    if not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    lib = C.CDLL(LIB_PATH)
    sig = {
        "oracle_box_nnodes": (_pd, [C.c_int] * 3),
        "oracle_box_nelements": (_pd, [C.c_int] * 4),
        "oracle_box_mesh": (None, [C.c_int] * 3 + [C.c_double] * 3 + [C.c_int, _i32p, _f32p]),
        "oracle_sshex8_nxe": (C.c_int, [C.c_int]),
        "oracle_sshex8_refine": (_pd, [C.c_int, _pd, _pd, _i32p, _f32p, C.c_void_p, C.c_void_p]),
        "oracle_sshex8_refine_ex": (_pd, [C.c_int, C.c_int, _pd, _pd, _i32p, _f32p, C.c_void_p, C.c_void_p]),
        "oracle_sshex8_hierarchical_levels": (C.c_int, [C.c_int, C.c_void_p]),
        "oracle_sshex8_derefine": (None, [C.c_int, C.c_int, _pd, _i32p, _i32p]),
        "oracle_sshex8_to_hex8": (None, [C.c_int, _pd, _i32p, _i32p]),
        "oracle_node_incidence": (None, [C.c_int, _pd, _pd, _i32p, _u16p]),
        "oracle_max_node": (_pd, [C.c_int, _pd, _i32p]),
        "oracle_tet4_fff": (None, [_pd, _i32p, _pd, _f32p, _f32p]),
        "oracle_hex8_fff": (None, [_pd, _i32p, _NullableI32, _pd, _f32p, _f32p]),
        "oracle_hex8_adjugate_det": (None, [_pd, _i32p, _NullableI32, _pd, _f32p, _f32p, _f32p]),
        "oracle_hex8_laplacian_matrix": (None, [_f64p, _f64p]),
        "oracle_tet4_laplacian_matrix": (None, [_f64p, _f64p]),
        "oracle_hex8_elasticity_matrix": (None, [C.c_double, C.c_double, _f64p, C.c_double, _f64p]),
        "oracle_tet4_laplacian_apply": (None, [_pd, _i32p, _f32p, _f64p, _f64p]),
        "oracle_tet4_laplacian_diag": (None, [_pd, _i32p, _f32p, _f64p]),
        "oracle_hex8_laplacian_apply": (None, [_pd, _i32p, _f32p, _NullableF32, _f64p, _f64p]),
        "oracle_hex8_laplacian_diag": (None, [_pd, _i32p, _f32p, _NullableF32, _f64p]),
        "oracle_sshex8_laplacian_apply": (None, [C.c_int, _pd, _i32p, _f32p, _NullableF32, _f64p, _f64p]),
        "oracle_sshex8_laplacian_stencil_apply": (None, [C.c_int, _pd, _i32p, _f32p, _NullableF32, _f64p, _f64p]),
        "oracle_sshex8_laplacian_diag": (None, [C.c_int, _pd, _i32p, _f32p, _NullableF32, _f64p]),
        "oracle_sshex8_em_apply": (None, [C.c_int, _pd, _i32p, C.c_int, _f64p, _f64p, _f64p]),
        "oracle_sshex8_laplacian_element_matrices": (None, [C.c_int, _pd, _f32p, _NullableF32, _f64p]),
        "oracle_sshex8_elasticity_element_matrices": (None, [C.c_int, _pd, _f32p, _f32p, C.c_double, C.c_double,
                                                             _NullableF32, _f64p]),
        "oracle_hex8_elasticity_apply": (None, [_pd, _i32p, _f32p, _f32p, C.c_double, C.c_double, _NullableF32,
                                                _f64p, _f64p]),
        "oracle_hex8_elasticity_diag": (None, [_pd, _i32p, _f32p, _f32p, C.c_double, C.c_double, _NullableF32,
                                               _f64p]),
        "oracle_sshex8_elasticity_apply": (None, [C.c_int, _pd, _i32p, _f32p, _f32p, C.c_double, C.c_double,
                                                  _NullableF32, _f64p, _f64p]),
        "oracle_sshex8_elasticity_diag": (None, [C.c_int, _pd, _i32p, _f32p, _f32p, C.c_double, C.c_double,
                                                 _NullableF32, _f64p]),
        "oracle_sshex8_elasticity_block_diag": (None, [C.c_int, _pd, _i32p, _f32p, _f32p, C.c_double,
                                                       C.c_double, _NullableF32, _f64p]),
        "oracle_sshex8_prolongate": (None, [_pd, C.c_int, _i32p, C.c_int, _i32p, C.c_int, _f64p, _f64p]),
        "oracle_sshex8_restrict": (None, [_pd, C.c_int, _i32p, _u16p, C.c_int, _i32p, C.c_int, _f64p, _f64p]),
        "oracle_sshex8_hierarchical_prolongation": (None, [C.c_int, _pd, _i32p, C.c_int, _f64p, _f64p]),
        "oracle_sshex8_hierarchical_restriction": (None, [C.c_int, _pd, _i32p, _u16p, C.c_int, _f64p, _f64p]),
        "oracle_constraint_nodes_to_value_vec": (None, [_pd, _i32p, C.c_int, C.c_int, C.c_double, _f64p]),
        "oracle_constraint_nodes_copy_vec": (None, [_pd, _i32p, C.c_int, C.c_int, _f64p, _f64p]),
        "oracle_num_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()

HEX8_CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def lidx(L, x, y, z):
    return (z * (L + 1) + y) * (L + 1) + x


def sshex8_corner_rows(L):
    return np.array([lidx(L, cx * L, cy * L, cz * L) for (cx, cy, cz) in HEX8_CORNERS], dtype=np.int32)


def box_mesh(nx, ny, nz, w=1.0, h=1.0, t=1.0, tet4=False):
    nn = lib.oracle_box_nnodes(nx, ny, nz)
    ne = lib.oracle_box_nelements(nx, ny, nz, int(tet4))
    el = np.zeros((4 if tet4 else 8, ne), np.int32)
    pts = np.zeros((3, nn), np.float32)
    lib.oracle_box_mesh(nx, ny, nz, w, h, t, int(tet4), el, pts)
    return el, pts


def sshex8_refine(L, macro_el, macro_pts, hierarchical=True):
    ne = macro_el.shape[1]
    nn = macro_pts.shape[1]
    macro_el = np.ascontiguousarray(macro_el, np.int32)
    macro_pts = np.ascontiguousarray(macro_pts, np.float32)
    h = int(bool(hierarchical))
    ntot = lib.oracle_sshex8_refine_ex(L, h, ne, nn, macro_el, macro_pts, None, None)
    el = np.zeros(((L + 1) ** 3, ne), np.int32)
    pts = np.zeros((3, ntot), np.float32)
    lib.oracle_sshex8_refine_ex(L, h, ne, nn, macro_el, macro_pts, el.ctypes.data, pts.ctypes.data)
    return el, pts


def hierarchical_levels(L):
    """sshex8_hierarchical_mesh_levels, coarsest first ([1, 2, 4, 8] for 8)."""
    buf = (C.c_int * 32)()
    n = lib.oracle_sshex8_hierarchical_levels(L, buf)
    return [buf[i] for i in range(n)]


def derefine(Lf, Lc, el):
    out = np.zeros(((Lc + 1) ** 3, el.shape[1]), np.int32)
    lib.oracle_sshex8_derefine(Lf, Lc, el.shape[1], np.ascontiguousarray(el), out)
    return out


def to_hex8(L, el):
    out = np.zeros((8, el.shape[1]), np.int32)
    lib.oracle_sshex8_to_hex8(L, el.shape[1], np.ascontiguousarray(el), out)
    return out


def incidence(el, nnodes):
    c = np.zeros(nnodes, np.uint16)
    lib.oracle_node_incidence(el.shape[0], el.shape[1], nnodes, np.ascontiguousarray(el), c)
    return c


def tet4_fff(el, pts):
    el, pts = np.ascontiguousarray(el, np.int32), np.ascontiguousarray(pts, np.float32)
    fff = np.zeros((el.shape[1], 6), np.float32)
    lib.oracle_tet4_fff(el.shape[1], el, pts.shape[1], pts, fff)
    return fff


def hex8_fff(el, pts, corners=None):
    el, pts = np.ascontiguousarray(el, np.int32), np.ascontiguousarray(pts, np.float32)
    fff = np.zeros((el.shape[1], 6), np.float32)
    lib.oracle_hex8_fff(el.shape[1], el, corners, pts.shape[1], pts, fff)
    return fff


def hex8_adj_det(el, pts, corners=None):
    el, pts = np.ascontiguousarray(el, np.int32), np.ascontiguousarray(pts, np.float32)
    adj = np.zeros((el.shape[1], 9), np.float32)
    det = np.zeros(el.shape[1], np.float32)
    lib.oracle_hex8_adjugate_det(el.shape[1], el, corners, pts.shape[1], pts, adj, det)
    return adj, det


def hex8_laplacian_matrix(fff):
    K = np.zeros(64)
    lib.oracle_hex8_laplacian_matrix(np.ascontiguousarray(fff, np.float64), K)
    return K.reshape(8, 8)


def tet4_laplacian_matrix(fff):
    K = np.zeros(16)
    lib.oracle_tet4_laplacian_matrix(np.ascontiguousarray(fff, np.float64), K)
    return K.reshape(4, 4)


def hex8_elasticity_matrix(mu, lam, adj, det):
    K = np.zeros(576)
    lib.oracle_hex8_elasticity_matrix(mu, lam, np.ascontiguousarray(adj, np.float64), det, K)
    return K.reshape(24, 24)


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
