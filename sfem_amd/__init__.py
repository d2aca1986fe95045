"""Python view of the sfem_amd C ABI (include/sfem_amd.h).

Mirrors the reference's operator/solver interface (pysfem, python/bindings/pysfem.cpp:585-719):
Operator.apply accumulates y += A x, solvers take x as initial guess and output.  Device
vectors are torch tensors on the GPU (torch is only the allocator here); every computation
runs in the library's own HIP kernels.  There is no CPU fallback: importing this module
raises if the compiled library is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SFEM_AMD_LIBRARY selects another build of the same library (e.g. a debug build)
LIB_PATH = os.environ.get("SFEM_AMD_LIBRARY") or os.path.join(_HERE, "lib", "libsfem_amd.so")

TET4, HEX8, SSHEX8 = 4, 8, 108
FLOAT64, FLOAT32 = 0, 1
SUCCESS, FAILURE = 0, 1

_vp, _pd, _i, _d = C.c_void_p, C.c_ssize_t, C.c_int, C.c_double

# name: (restype, argtypes) -- exactly the declarations of include/sfem_amd.h
SIGNATURES = {
    "sfem_amd_last_error": (C.c_char_p, []),
    "sfem_amd_device_count": (_i, []),
    "sfem_amd_device_synchronize": (_i, []),
    "sfem_amd_malloc": (_vp, [C.c_size_t]),
    "sfem_amd_free": (_i, [_vp]),
    "sfem_amd_memcpy_h2d": (_i, [_vp, _vp, C.c_size_t]),
    "sfem_amd_memcpy_d2h": (_i, [_vp, _vp, C.c_size_t]),
    "sfem_amd_memset": (_i, [_vp, _i, C.c_size_t, _vp]),
    "sfem_amd_box_nnodes": (_pd, [_i, _i, _i]),
    "sfem_amd_box_nelements": (_pd, [_i, _i, _i, _i]),
    "sfem_amd_box_mesh": (_i, [_i, _i, _i, _d, _d, _d, _i, _vp, _vp]),
    "sfem_amd_sshex8_refine": (_pd, [_i, _pd, _pd, _vp, _vp, _vp, _vp]),
    "sfem_amd_sshex8_refine_ex": (_pd, [_i, _i, _pd, _pd, _vp, _vp, _vp, _vp]),
    "sfem_amd_fff": (_i, [_i, _i, _pd, _pd, _vp, _pd, _vp, _vp]),
    "sfem_amd_adjugate_det": (_i, [_i, _i, _pd, _pd, _vp, _pd, _vp, _vp, _vp]),
    "sfem_amd_mesh_create": (_i, [_i, _i, _pd, _pd, _vp, _pd, C.POINTER(_vp)]),
    "sfem_amd_mesh_derefine": (_i, [_vp, _i, C.POINTER(_vp)]),
    "sfem_amd_sshex8_refine_device": (_i, [_i, _pd, _pd, _vp, C.POINTER(_vp), _vp]),
    "sfem_amd_sshex8_refine_points_device": (_i, [_vp, _pd, _vp, _vp, _vp, _vp]),
    "sfem_amd_mesh_fff": (_i, [_vp, _vp, _vp, _vp]),
    "sfem_amd_mesh_adjugate_det": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "sfem_amd_mesh_nnodes": (_pd, [_vp]),
    "sfem_amd_mesh_nelements": (_pd, [_vp]),
    "sfem_amd_mesh_level": (_i, [_vp]),
    "sfem_amd_mesh_element_type": (_i, [_vp]),
    "sfem_amd_mesh_download": (_i, [_vp, _vp]),
    "sfem_amd_mesh_destroy": (None, [_vp]),
    "sfem_amd_laplacian_create": (_i, [_vp, _vp, _vp, _i, C.POINTER(_vp)]),
    "sfem_amd_linear_elasticity_create": (_i, [_vp, _vp, _vp, _d, _d, _vp, _i, C.POINTER(_vp)]),
    "sfem_amd_op_apply": (_i, [_vp, _vp, _vp, _vp]),
    "sfem_amd_op_diag": (_i, [_vp, _vp, _vp]),
    "sfem_amd_op_block_diag_sym": (_i, [_vp, _vp, _vp]),
    "sfem_amd_op_rows": (_pd, [_vp]),
    "sfem_amd_op_cols": (_pd, [_vp]),
    "sfem_amd_op_real_type": (_i, [_vp]),
    "sfem_amd_op_destroy": (None, [_vp]),
    "sfem_amd_constrained_op_create": (_i, [_vp, _pd, _vp, _i, _i, _vp, C.POINTER(_vp)]),
    "sfem_amd_constrained_op_set_value": (_i, [_vp, _d, _vp, _vp]),
    "sfem_amd_prolongation_create": (_i, [_vp, _vp, _i, _i, _vp, C.POINTER(_vp)]),
    "sfem_amd_restriction_create": (_i, [_vp, _vp, _i, _i, _vp, C.POINTER(_vp)]),
    "sfem_amd_blas_axpby": (_i, [_i, _pd, _d, _vp, _d, _vp, _vp]),
    "sfem_amd_blas_zaxpby": (_i, [_i, _pd, _d, _vp, _d, _vp, _vp, _vp]),
    "sfem_amd_blas_scal": (_i, [_i, _pd, _d, _vp, _vp]),
    "sfem_amd_blas_values": (_i, [_i, _pd, _d, _vp, _vp]),
    "sfem_amd_blas_copy": (_i, [_i, _pd, _vp, _vp, _vp]),
    "sfem_amd_blas_reciprocal": (_i, [_i, _pd, _d, _vp, _vp]),
    "sfem_amd_blas_xypaz": (_i, [_i, _pd, _vp, _vp, _d, _vp, _vp]),
    "sfem_amd_blas_dot": (_d, [_i, _pd, _vp, _vp, _vp]),
    "sfem_amd_blas_norm2": (_d, [_i, _pd, _vp, _vp]),
    "sfem_amd_blas_dot_noalloc": (_i, [_i, _pd, _vp, _vp, _vp, _vp]),
    "sfem_amd_blas_cast": (_i, [_i, _pd, _vp, _i, _vp, _vp]),
    "sfem_amd_cg_update_x_r": (_d, [_i, _pd, _d, _vp, _vp, _vp, _vp, _i, _vp]),
    "sfem_amd_cg_update_x_r_dot": (_d, [_i, _pd, _d, _vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_double), _vp]),
    "sfem_amd_cheb3_update": (_i, [_i, _pd, _d, _d, _vp, _vp, _vp, _vp, _vp]),
    "sfem_amd_cheb3_update_diag": (_i, [_i, _pd, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sfem_amd_cheb3_start": (_i, [_i, _pd, _d, _vp, _vp, _i, _vp, _vp, _vp]),
    "sfem_amd_set_option": (_i, [C.c_char_p, _i]),
    "sfem_amd_cheb3_diag_create": (_i, [_vp, _vp, _i, _d, _d, _i, _d, _i, _vp, C.POINTER(_vp)]),
    "sfem_amd_cg_create": (_i, [_vp, _vp, _i, _d, _d, _i, C.POINTER(_vp)]),
    "sfem_amd_cheb3_create": (_i, [_vp, _i, _d, _d, _i, _d, _i, _vp, C.POINTER(_vp)]),
    "sfem_amd_cheb3_eig_max": (_d, [_vp]),
    "sfem_amd_cheb3_lanczos_create": (_i, [_vp, _vp, _i, _d, _d, _i, _d, _i, _vp, C.POINTER(_vp)]),
    "sfem_amd_cheb3_eigen_iterations": (_i, [_vp]),
    "sfem_amd_jacobi_create": (_i, [_vp, _vp, _d, _i, C.POINTER(_vp)]),
    "sfem_amd_inverse_diagonal_create": (_i, [_i, _pd, _vp, _d, C.POINTER(_vp)]),
    "sfem_amd_permutation_create": (_i, [_pd, _vp, _i, _i, _i, _vp, C.POINTER(_vp)]),
    "sfem_amd_chain_create": (_i, [_vp, _vp, C.POINTER(_vp)]),
    "sfem_amd_solver_bridge_create": (_i, [_vp, C.POINTER(_vp)]),
    "sfem_amd_mg_create": (_i, [_i, _i, _i, _d, C.POINTER(_vp)]),
    "sfem_amd_mg_add_level": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "sfem_amd_precision_bridge_create": (_i, [_vp, C.POINTER(_vp)]),
    "sfem_amd_op_iterations": (_i, [_vp]),
    "sfem_amd_cg_history": (_i, [_vp, _vp, _i]),
    "sfem_amd_op_last_status": (_i, [_vp]),
    "sfem_amd_op_benchmark": (_i, [_vp, _vp, _vp, _i, _i, C.c_size_t, _vp, C.POINTER(_d)]),
    "sfem_amd_element_matrix_op_create": (_i, [_vp, _vp, _i, _i, C.POINTER(_vp)]),
    "sfem_amd_cnn_create": (_i, [_i, _i, _vp, C.c_uint, C.POINTER(_vp)]),
    "sfem_amd_cnn_create_channels": (_i, [_i, _i, _i, _vp, C.c_uint, C.POINTER(_vp)]),
    "sfem_amd_cnn_channels": (_i, [_vp]),
    "sfem_amd_cnn_num_params": (_i, [_vp]),
    "sfem_amd_cnn_get_weights": (_i, [_vp, _vp]),
    "sfem_amd_cnn_set_weights": (_i, [_vp, _vp]),
    "sfem_amd_cnn_destroy": (None, [_vp]),
    "sfem_amd_cnn_forward": (_i, [_vp, _i, _pd, _vp, _vp, _i, _vp]),
    "sfem_amd_cnn_loss_grad": (_i, [_vp, _i, _pd, _vp, _vp, C.POINTER(_d), _vp, _vp]),
    "sfem_amd_cnn_adam_step": (_i, [_vp, _d, _vp]),
    "sfem_amd_gather_elements": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "sfem_amd_learned_smoother_create": (_i, [_vp, _vp, _vp, _vp, _i, _i, _d, _i, C.POINTER(_vp)]),
    "sfem_amd_dense_op_create": (_i, [_i, _pd, _vp, C.POINTER(_vp)]),
    "sfem_amd_block_jacobi_create": (_i, [_vp, _vp, _vp, _d, _i, C.POINTER(_vp)]),
    "sfem_amd_graph_op_create": (_i, [_vp, C.POINTER(_vp)]),
    "sfem_amd_graph_op_stats": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "sfem_amd_switch_create": (_i, [_vp, _vp, C.POINTER(_vp)]),
    "sfem_amd_switch_select": (_i, [_vp, _i]),
    "sfem_amd_switch_state": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "sfem_amd_cg_add_fallback": (_i, [_vp, _vp, _d]),
}


def _preload_shared_hip_runtime():
    """torch-ROCm wheels bundle a HIP runtime with the system runtime's SONAME
    (libamdhip64.so.7).  Whichever loads first serves the whole process, and torch cannot
    initialise on a runtime another library brought in.  Loading torch's copy first (without
    importing torch) gives the library and torch one shared runtime in every import order."""
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.submodule_search_locations:
        return
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(path):
        C.CDLL(path, mode=C.RTLD_GLOBAL)


def _load():
    _preload_shared_hip_runtime()
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"sfem_amd: compiled library not found at {LIB_PATH}; build it with "
            "`make -C sfem_amd` or __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


class SfemError(RuntimeError):
    pass


def _check(ret, what):
    if ret != SUCCESS:
        raise SfemError(f"{what}: {lib.sfem_amd_last_error().decode()}")


def _np_ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _host(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- host mesh setup
def box_mesh(nx, ny, nz, w=1.0, h=1.0, t=1.0, element_type=HEX8):
    """box_mesh.py:479-700 layout: (elements [nv][E] int32, points [3][N] float32)."""
    nn = lib.sfem_amd_box_nnodes(nx, ny, nz)
    ne = lib.sfem_amd_box_nelements(nx, ny, nz, element_type)
    el = np.zeros((4 if element_type == TET4 else 8, ne), np.int32)
    pts = np.zeros((3, nn), np.float32)
    _check(lib.sfem_amd_box_mesh(nx, ny, nz, w, h, t, element_type, _np_ptr(el), _np_ptr(pts)), "box_mesh")
    return el, pts


def sshex8_refine(level, macro_elements, macro_points, hierarchical=True):
    """smesh::to_semistructured(level, mesh, hierarchical, false): connectivity [(L+1)^3][E] and
    points [3][N] (hierarchical numbering by default: every derefined level is a prefix)."""
    me = _host(macro_elements, np.int32)
    mp = _host(macro_points, np.float32)
    h = int(bool(hierarchical))
    n = lib.sfem_amd_sshex8_refine_ex(level, h, me.shape[1], mp.shape[1], _np_ptr(me), _np_ptr(mp), None, None)
    if n < 0:
        raise SfemError(lib.sfem_amd_last_error().decode())
    el = np.zeros(((level + 1) ** 3, me.shape[1]), np.int32)
    pts = np.zeros((3, n), np.float32)
    n2 = lib.sfem_amd_sshex8_refine_ex(level, h, me.shape[1], mp.shape[1], _np_ptr(me), _np_ptr(mp), _np_ptr(el),
                                       _np_ptr(pts))
    if n2 != n:
        raise SfemError(lib.sfem_amd_last_error().decode())
    return el, pts


def fff(element_type, level, elements, points):
    el = _host(elements, np.int32)
    p = _host(points, np.float32)
    out = np.zeros((el.shape[1], 6), np.float32)
    _check(lib.sfem_amd_fff(element_type, level, el.shape[1], el.shape[1], _np_ptr(el), p.shape[1], _np_ptr(p),
                            _np_ptr(out)), "fff")
    return out


def adjugate_det(element_type, level, elements, points):
    el = _host(elements, np.int32)
    p = _host(points, np.float32)
    adj = np.zeros((el.shape[1], 9), np.float32)
    det = np.zeros(el.shape[1], np.float32)
    _check(lib.sfem_amd_adjugate_det(element_type, level, el.shape[1], el.shape[1], _np_ptr(el), p.shape[1],
                                     _np_ptr(p), _np_ptr(adj), _np_ptr(det)), "adjugate_det")
    return adj, det


# ---------------------------------------------------------------- device memory
class DeviceArray:
    """A flat device vector owned by the library's allocator (no torch needed)."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = lib.sfem_amd_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise SfemError("sfem_amd_malloc: " + lib.sfem_amd_last_error().decode())
        self.zero()

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype).ravel()
        d = cls(a.size, a.dtype)
        if a.size:
            _check(lib.sfem_amd_memcpy_h2d(d.ptr, a.ctypes.data_as(_vp), d.nbytes), "memcpy_h2d")
        return d

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        if self.n:
            _check(lib.sfem_amd_memcpy_d2h(out.ctypes.data_as(_vp), self.ptr, self.nbytes), "memcpy_d2h")
        return out

    def zero(self, stream=None):
        _check(lib.sfem_amd_memset(self.ptr, 0, self.nbytes, _stream_ptr(stream)), "memset")
        return self

    def _real_type(self):
        if self.dtype not in (np.float64, np.float32):
            raise SfemError("DeviceArray: a float64 or float32 vector required")
        return FLOAT64 if self.dtype == np.float64 else FLOAT32

    def fill(self, value, stream=None):
        """Every entry = value (on the device)."""
        _check(lib.sfem_amd_blas_values(self._real_type(), self.n, float(value), self.ptr, _stream_ptr(stream)),
               "blas_values")
        return self

    def reciprocal(self, alpha=1.0, stream=None):
        """x = alpha / x entrywise (on the device)."""
        _check(lib.sfem_amd_blas_reciprocal(self._real_type(), self.n, float(alpha), self.ptr, _stream_ptr(stream)),
               "blas_reciprocal")
        return self

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        assert a.size == self.n
        _check(lib.sfem_amd_memcpy_h2d(self.ptr, a.ctypes.data_as(_vp), self.nbytes), "memcpy_h2d")
        return self

    def numel(self):
        return self.n

    def element_size(self):
        return self.dtype.itemsize

    def __del__(self):
        if getattr(self, "ptr", None) and lib is not None:
            lib.sfem_amd_free(self.ptr)
            self.ptr = None


def _dptr(t):
    """Device pointer of a DeviceArray or of a contiguous GPU torch tensor."""
    if t is None:
        return None
    if isinstance(t, DeviceArray):
        return C.c_void_p(t.ptr)
    if not (hasattr(t, "is_cuda") and t.is_cuda and t.is_contiguous()):
        raise SfemError("expected a DeviceArray or a contiguous GPU tensor")
    return C.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    if stream is None:
        return None
    return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


class Mesh:
    """Device-resident connectivity (sfem_amd_mesh)."""

    def __init__(self, element_type, level, elements, nnodes, _handle=None):
        self._keep = []
        if _handle is not None:
            self.h = _handle
        else:
            el = _host(elements, np.int32)
            h = _vp()
            _check(lib.sfem_amd_mesh_create(element_type, level, el.shape[1], el.shape[1], _np_ptr(el), nnodes,
                                            C.byref(h)), "mesh_create")
            self.h = h
        self.element_type = lib.sfem_amd_mesh_element_type(self.h)
        self.level = lib.sfem_amd_mesh_level(self.h)
        self.nnodes = lib.sfem_amd_mesh_nnodes(self.h)
        self.nelements = lib.sfem_amd_mesh_nelements(self.h)

    def derefine(self, level):
        h = _vp()
        _check(lib.sfem_amd_mesh_derefine(self.h, level, C.byref(h)), "mesh_derefine")
        m = Mesh(None, None, None, None, _handle=h)
        m._keep.append(self)
        return m

    def download(self):
        nv = 8 if self.element_type == HEX8 else 4 if self.element_type == TET4 else (self.level + 1) ** 3
        out = np.zeros((nv, self.nelements), np.int32)
        _check(lib.sfem_amd_mesh_download(self.h, _np_ptr(out)), "mesh_download")
        return out

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:
            lib.sfem_amd_mesh_destroy(self.h)
            self.h = None


class Operator:
    """An sfem_amd_op handle: FE operator, transfer, constrained operator or solver."""

    def __init__(self, handle, keep=()):
        self.h = handle
        self._keep = list(keep)

    @property
    def rows(self):
        return lib.sfem_amd_op_rows(self.h)

    @property
    def cols(self):
        return lib.sfem_amd_op_cols(self.h)

    @property
    def real_type(self):
        return lib.sfem_amd_op_real_type(self.h)

    def _check_vec(self, t, n):
        size = 8 if self.real_type == FLOAT64 else 4
        if t.element_size() != size or t.numel() != n:
            raise SfemError(f"vector must have {size}-byte reals and {n} entries "
                            f"(got {t.element_size()}-byte, {t.numel()})")

    def apply(self, x, y, stream=None):
        self._check_vec(x, self.cols)
        self._check_vec(y, self.rows)
        _check(lib.sfem_amd_op_apply(self.h, _dptr(x), _dptr(y), _stream_ptr(stream)), "op_apply")
        return y

    def diag(self, d, stream=None):
        self._check_vec(d, self.rows)
        _check(lib.sfem_amd_op_diag(self.h, _dptr(d), _stream_ptr(stream)), "op_diag")
        return d

    def block_diag_sym(self, d, stream=None):
        _check(lib.sfem_amd_op_block_diag_sym(self.h, _dptr(d), _stream_ptr(stream)), "op_block_diag_sym")
        return d

    @property
    def iterations(self):
        return lib.sfem_amd_op_iterations(self.h)

    @property
    def last_status(self):
        return lib.sfem_amd_op_last_status(self.h)

    def history(self):
        n = lib.sfem_amd_cg_history(self.h, None, 0)
        out = np.zeros(max(n, 0))
        if n > 0:
            lib.sfem_amd_cg_history(self.h, _np_ptr(out), n)
        return out

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:
            lib.sfem_amd_op_destroy(self.h)
            self.h = None


def _new_op(fn, *args, keep=()):
    h = _vp()
    _check(getattr(lib, fn)(*args, C.byref(h)), fn)
    return Operator(h, keep)


def _geom(a):
    """(pointer, keep-alive) of a geometry array: a DeviceArray is passed as device memory."""
    if isinstance(a, DeviceArray):
        return _vp(a.ptr), a
    h = _host(a, np.float32)
    return _np_ptr(h), h


def laplacian(mesh, fff_aos, coeff=None, real_type=FLOAT64):
    """fff_aos [E][6] (numpy, host) or a DeviceArray (e.g. from mesh_fff); coeff likewise."""
    f, fk = _geom(fff_aos)
    c, ck = _geom(coeff)
    return _new_op("sfem_amd_laplacian_create", mesh.h, f, c, real_type, keep=[mesh])


def linear_elasticity(mesh, adjugate, det, mu, lam, coeff=None, real_type=FLOAT64):
    a, ak = _geom(adjugate)
    d, dk = _geom(det)
    c, ck = _geom(coeff)
    return _new_op("sfem_amd_linear_elasticity_create", mesh.h, a, d, mu, lam, c, real_type, keep=[mesh])


def sshex8_refine_device(level, macro_elements, macro_points=None, nnodes=None, stream=None):
    """Refinement on the device: macro_elements [8][E] (numpy or DeviceArray int32, SoA) and
    optionally macro_points [3][N] (numpy or DeviceArray float32).  Returns (Mesh, points) with
    points a DeviceArray [3 * nnodes] (or None); ids identical to sshex8_refine."""
    if isinstance(macro_elements, DeviceArray):
        me = macro_elements
        ne = me.n // 8
    else:
        h = np.ascontiguousarray(macro_elements, np.int32)
        ne = h.shape[1]
        me = DeviceArray.from_numpy(h)
    mp = None
    if macro_points is not None:
        mp = macro_points if isinstance(macro_points, DeviceArray) else \
            DeviceArray.from_numpy(np.ascontiguousarray(macro_points, np.float32))
        nnodes = mp.n // 3
    if nnodes is None:
        raise ValueError("sshex8_refine_device: nnodes (or macro_points) required")
    h = _vp()
    _check(lib.sfem_amd_sshex8_refine_device(level, ne, nnodes, me.ptr, C.byref(h), _stream_ptr(stream)),
           "sshex8_refine_device")
    mesh = Mesh(None, None, None, None, _handle=h)
    pts = None
    if mp is not None:
        pts = DeviceArray(3 * mesh.nnodes, np.float32)
        _check(lib.sfem_amd_sshex8_refine_points_device(mesh.h, nnodes, me.ptr, mp.ptr, pts.ptr,
                                                         _stream_ptr(stream)), "sshex8_refine_points_device")
    return mesh, pts


def mesh_fff(mesh, points, stream=None):
    """FFF [E][6] of a device mesh from device points [3][nnodes] -> DeviceArray float32."""
    out = DeviceArray(6 * mesh.nelements, np.float32)
    _check(lib.sfem_amd_mesh_fff(mesh.h, points.ptr, out.ptr, _stream_ptr(stream)), "mesh_fff")
    return out


def mesh_adjugate_det(mesh, points, stream=None):
    adj = DeviceArray(9 * mesh.nelements, np.float32)
    det = DeviceArray(mesh.nelements, np.float32)
    _check(lib.sfem_amd_mesh_adjugate_det(mesh.h, points.ptr, adj.ptr, det.ptr, _stream_ptr(stream)),
           "mesh_adjugate_det")
    return adj, det


def element_matrix_op(mesh, matrices, block_size=1, real_type=FLOAT64):
    """ss:em operator: stored per-macro sub-element matrices ([E][64] or [E][576], FP64)."""
    m = _host(matrices, np.float64)
    return _new_op("sfem_amd_element_matrix_op_create", mesh.h, _np_ptr(m), block_size, real_type, keep=[mesh])


def constrained(op, nodes, block_size=1, components=None):
    n = _host(nodes, np.int32)
    comps = None if components is None else np.ascontiguousarray(components, np.int32)
    nc = 0 if comps is None else len(comps)
    return _new_op("sfem_amd_constrained_op_create", op.h, len(n), _np_ptr(n), block_size, nc, _np_ptr(comps),
                   keep=[op])


def set_constrained_value(cop, value, x, stream=None):
    _check(lib.sfem_amd_constrained_op_set_value(cop.h, value, _dptr(x), _stream_ptr(stream)), "set_value")


def prolongation(coarse, fine, block_size=1, real_type=FLOAT64, constraints=None):
    return _new_op("sfem_amd_prolongation_create", coarse.h, fine.h, block_size, real_type,
                   constraints.h if constraints else None, keep=[coarse, fine, constraints])


def restriction(fine, coarse, block_size=1, real_type=FLOAT64, constraints=None):
    return _new_op("sfem_amd_restriction_create", fine.h, coarse.h, block_size, real_type,
                   constraints.h if constraints else None, keep=[coarse, fine, constraints])


def cg(A, precond=None, flexible=False, rtol=1e-10, atol=1e-16, max_it=10000, stop_on_residual=False):
    """mode: flexible -> 1 (flexible PCG, true-residual stop), stop_on_residual -> 2 (fixed-preconditioner
    PCG stopping on ||r||/||b||), else 0 (the reference's PCG)."""
    mode = 1 if flexible else (2 if stop_on_residual else 0)
    return _new_op("sfem_amd_cg_create", A.h, precond.h if precond else None, mode, rtol, atol, max_it,
                   keep=[A, precond])


def cheb3(A, max_it=3, scale_eig_max=1.0, scale_eig_min=0.06, eigen_solver_max_it=1000, eigen_solver_tol=1e-6,
          initial_guess_zero=False, init_vector=None, inv_diag=None, lanczos_steps=0):
    """Chebyshev3 smoother; with inv_diag (device D^-1) it runs on D^-1 A (an extension).  eig_max
    from the power method (the reference's), or with lanczos_steps > 0 from at most that many
    Lanczos steps (an extension; eigen_solver_tol is then a relative tolerance)."""
    if lanczos_steps > 0:
        return _new_op("sfem_amd_cheb3_lanczos_create", A.h, _dptr(inv_diag) if inv_diag is not None else None,
                       max_it, scale_eig_max, scale_eig_min, lanczos_steps, eigen_solver_tol,
                       int(initial_guess_zero), _dptr(init_vector), keep=[A])
    if inv_diag is not None:
        return _new_op("sfem_amd_cheb3_diag_create", A.h, _dptr(inv_diag), max_it, scale_eig_max, scale_eig_min,
                       eigen_solver_max_it, eigen_solver_tol, int(initial_guess_zero), _dptr(init_vector), keep=[A])
    return _new_op("sfem_amd_cheb3_create", A.h, max_it, scale_eig_max, scale_eig_min, eigen_solver_max_it,
                   eigen_solver_tol, int(initial_guess_zero), _dptr(init_vector), keep=[A])


def cheb3_eig_max(op):
    return lib.sfem_amd_cheb3_eig_max(op.h)


def cheb3_eigen_iterations(op):
    """Iterations of the smoother's eigenvalue estimate (power method or Lanczos)."""
    return lib.sfem_amd_cheb3_eigen_iterations(op.h)


def jacobi(A, diag, omega=1.0, steps=3):
    return _new_op("sfem_amd_jacobi_create", A.h, _dptr(diag), omega, steps, keep=[A])


def block_jacobi(A, block_diag_sym, dof_mask=None, omega=1.0 / 3, steps=3):
    """Block-Jacobi stationary smoother on 3x3 blocks (block_diag_sym: device [n/3][6];
    dof_mask: device uint8 [n], 1 = constrained)."""
    return _new_op("sfem_amd_block_jacobi_create", A.h, _dptr(block_diag_sym), _dptr(dof_mask), omega, steps,
                   keep=[A])


def permutation(perm, block_size=1, real_type=FLOAT64, transpose=False, constraints=None):
    p = _host(perm, np.int32)
    return _new_op("sfem_amd_permutation_create", len(p), _np_ptr(p), block_size, real_type, int(transpose),
                   constraints.h if constraints else None, keep=[constraints])


def chain(A, B):
    """y += B (A x)"""
    return _new_op("sfem_amd_chain_create", A.h, B.h, keep=[A, B])


def inverse_diagonal(diag, omega=1.0):
    rt = FLOAT64 if diag.element_size() == 8 else FLOAT32
    return _new_op("sfem_amd_inverse_diagonal_create", rt, diag.numel(), _dptr(diag), omega)


class Multigrid(Operator):
    """sfem::Multigrid (algebra/sfem_Multigrid.hpp:27-347); level 0 is the finest."""

    def __init__(self, real_type=FLOAT64, cycle_type=1, max_it=10, atol=1e-10):
        h = _vp()
        _check(lib.sfem_amd_mg_create(real_type, cycle_type, max_it, atol, C.byref(h)), "mg_create")
        super().__init__(h)

    def add_level(self, op, smoother, prolongation=None, restriction=None):
        _check(lib.sfem_amd_mg_add_level(self.h, op.h, smoother.h, prolongation.h if prolongation else None,
                                         restriction.h if restriction else None), "mg_add_level")
        self._keep += [op, smoother, prolongation, restriction]


def precision_bridge(op):
    return _new_op("sfem_amd_precision_bridge_create", op.h, keep=[op])


def solver_bridge(solver):
    return _new_op("sfem_amd_solver_bridge_create", solver.h, keep=[solver])


class CNN:
    """The learned smoother's network (sfem_amd_cnn): 3x3x3 convolutions without biases,
    channels -> hidden -> ... -> channels with ReLU between, f(z) = g(z) - g(-z) per macro element;
    channels 1 (scalar problems) or 3 (linear elasticity, one channel per component)."""

    def __init__(self, hidden=16, layers=4, weights=None, seed=0, channels=1):
        h = _vp()
        w = None if weights is None else _host(weights, np.float32)
        _check(lib.sfem_amd_cnn_create_channels(channels, hidden, layers, None if w is None else _np_ptr(w), seed,
                                                C.byref(h)), "cnn_create")
        self.h = h
        self.hidden, self.layers, self.channels = hidden, layers, channels

    @property
    def num_params(self):
        return lib.sfem_amd_cnn_num_params(self.h)

    def get_weights(self):
        w = np.zeros(self.num_params, np.float32)
        _check(lib.sfem_amd_cnn_get_weights(self.h, _np_ptr(w)), "cnn_get_weights")
        return w

    def set_weights(self, w):
        w = _host(w, np.float32)
        if w.size != self.num_params:
            raise SfemError(f"cnn: {self.num_params} weights expected, got {w.size}")
        _check(lib.sfem_amd_cnn_set_weights(self.h, _np_ptr(w)), "cnn_set_weights")

    def forward(self, level, z, out, mode=0, stream=None):
        n = (level + 1) ** 3 * self.channels
        if z.numel() % n or out.numel() != z.numel() or z.element_size() != 4 or out.element_size() != 4:
            raise SfemError("cnn_forward: float32 [E][channels][(L+1)^3] input and output expected")
        _check(lib.sfem_amd_cnn_forward(self.h, level, z.numel() // n, _dptr(z), _dptr(out), mode,
                                        _stream_ptr(stream)), "cnn_forward")
        return out

    def loss_grad(self, level, z, target, want_grad=True, stream=None):
        n = (level + 1) ** 3 * self.channels
        if z.numel() % n or target.numel() != z.numel():
            raise SfemError("cnn_loss_grad: float32 [E][channels][(L+1)^3] input and target expected")
        loss = C.c_double()
        g = np.zeros(self.num_params, np.float32) if want_grad else None
        _check(lib.sfem_amd_cnn_loss_grad(self.h, level, z.numel() // n, _dptr(z), _dptr(target), C.byref(loss),
                                          None if g is None else _np_ptr(g), _stream_ptr(stream)), "cnn_loss_grad")
        return loss.value, g

    def adam_step(self, lr, stream=None):
        _check(lib.sfem_amd_cnn_adam_step(self.h, lr, _stream_ptr(stream)), "cnn_adam_step")

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:
            lib.sfem_amd_cnn_destroy(self.h)
            self.h = None


def gather_elements(mesh, vec, out, block_size=1, component=0, scale=None, stream=None):
    """out[e][v] = vec[idx[e][v]*bs + comp] (* scale[...]) as float32 (sshex8 meshes); component -1:
    all components as channels, out [E][bs][(L+1)^3]."""
    rt = FLOAT64 if vec.element_size() == 8 else FLOAT32
    _check(lib.sfem_amd_gather_elements(mesh.h, rt, block_size, component, _dptr(vec),
                                        None if scale is None else _dptr(scale), _dptr(out), _stream_ptr(stream)),
           "gather_elements")
    return out


def learned_smoother(A, mesh, cnn, diag, block_size=1, steps=1, damping=1.0, mode=1):
    return _new_op("sfem_amd_learned_smoother_create", A.h, mesh.h, cnn.h, _dptr(diag), block_size, steps,
                   damping, mode, keep=[A, mesh, cnn])


def dense_op(matrix, real_type=FLOAT64):
    """y += M x with a dense FP64 matrix (e.g. the inverse of a coarsest-level operator)."""
    m = _host(matrix, np.float64)
    if m.ndim != 2 or m.shape[0] != m.shape[1]:
        raise SfemError("dense_op: square matrix expected")
    return _new_op("sfem_amd_dense_op_create", real_type, m.shape[0], _np_ptr(m))


def graph_op(op):
    """Replays op's apply as a captured HIP graph (see sfem_amd_graph_op_create)."""
    return _new_op("sfem_amd_graph_op_create", op.h, keep=[op])


def graph_op_stats(op):
    c, r = C.c_int(), C.c_int()
    _check(lib.sfem_amd_graph_op_stats(op.h, C.byref(c), C.byref(r)), "graph_op_stats")
    return c.value, r.value


def switch(primary, fallback):
    return _new_op("sfem_amd_switch_create", primary.h, fallback.h, keep=[primary, fallback])


def switch_select(sw, which):
    _check(lib.sfem_amd_switch_select(sw.h, which), "switch_select")


def switch_state(sw):
    a, e = C.c_int(), C.c_int()
    _check(lib.sfem_amd_switch_state(sw.h, C.byref(a), C.byref(e)), "switch_state")
    return a.value, e.value


def cg_add_fallback(cg_op, sw, max_ratio=0.5):
    _check(lib.sfem_amd_cg_add_fallback(cg_op.h, sw.h, max_ratio), "cg_add_fallback")
    cg_op._keep.append(sw)


def benchmark(op, x, y, warmup=2, reps=10, flush_bytes=0, stream=None):
    """Total elapsed ms (HIP events) of `reps` applies; see sfem_amd_op_benchmark."""
    ms = C.c_double()
    _check(lib.sfem_amd_op_benchmark(op.h, _dptr(x), _dptr(y), warmup, reps, flush_bytes, _stream_ptr(stream),
                                     C.byref(ms)), "op_benchmark")
    return ms.value


def dot(x, y, stream=None):
    rt = FLOAT64 if x.element_size() == 8 else FLOAT32
    return lib.sfem_amd_blas_dot(rt, x.numel(), _dptr(x), _dptr(y), _stream_ptr(stream))


def norm2(x, stream=None):
    rt = FLOAT64 if x.element_size() == 8 else FLOAT32
    return lib.sfem_amd_blas_norm2(rt, x.numel(), _dptr(x), _stream_ptr(stream))


def set_option(name, value):
    """Library options, see sfem_amd_set_option ("fused_chebyshev", "atomic_skeleton": 1 or 0,
    default 0; "elasticity_form": 0 auto, 1 element form, 2 stencil form; "hex8_patches": 1 (default)
    or 0)."""
    _check(lib.sfem_amd_set_option(name.encode(), int(value)), "set_option")


def device_count():
    return lib.sfem_amd_device_count()


def synchronize():
    _check(lib.sfem_amd_device_synchronize(), "synchronize")
