"""Mesh and geometry setup on the device against the host path (which the oracle pins bit-exactly,
tests/test_capi_host.py): refined connectivity, node count, refined points and geometry caches
bit-exact (neither side contracts into FMA; shared nodes are written by their lowest element), and
operators built from device geometry equal to the host-built ones."""
import time

import numpy as np
import pytest

import oracle_lib as O

S = pytest.importorskip("sfem_amd")

pytestmark = pytest.mark.gpu


def permuted_box(n, seed):
    """A box of n^3 hex8 elements with its node ids randomly permuted (non-monotone ids exercise
    the first-occurrence ranking and the edge/face orientation rules) and slightly perturbed points."""
    el, pts = O.box_mesh(n + 1, n + 1, n + 1)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(pts.shape[1]).astype(np.int32)
    el = perm[el].astype(np.int32)
    p2 = np.empty_like(pts)
    p2[:, perm] = pts
    p2 = (p2 + 0.01 * rng.uniform(-1, 1, p2.shape)).astype(np.float32)
    return el, p2


@pytest.mark.parametrize("L,n,seed", [(1, 2, 0), (2, 3, 1), (3, 2, 2), (4, 3, 3), (6, 2, 4), (8, 2, 5), (16, 1, 6)])
def test_device_refine_matches_host(L, n, seed):
    mel, mpts = permuted_box(n, seed)
    el_h, pts_h = S.sshex8_refine(L, mel, mpts)
    mesh, pts_d = S.sshex8_refine_device(L, mel, mpts)
    assert mesh.nnodes == pts_h.shape[1]
    np.testing.assert_array_equal(mesh.download(), el_h)
    np.testing.assert_array_equal(pts_d.numpy().reshape(3, -1), pts_h)


def test_device_tet4_hex8_geometry_match_host():
    """tet4 FFF on the device follows the reference's float32 evaluation as the host path does."""
    rng = np.random.default_rng(21)
    for tet4 in (True, False):
        el, pts = O.box_mesh(6, 5, 7, tet4=tet4)
        pts = (pts + 0.03 * rng.uniform(-1, 1, pts.shape)).astype(np.float32)
        typ = S.TET4 if tet4 else S.HEX8
        mesh = S.Mesh(typ, 1, el, pts.shape[1])
        fff_d = S.mesh_fff(mesh, S.DeviceArray.from_numpy(pts.ravel()))
        np.testing.assert_array_equal(fff_d.numpy().reshape(-1, 6), S.fff(typ, 1, el, pts))
        if not tet4:
            adj_d, det_d = S.mesh_adjugate_det(mesh, S.DeviceArray.from_numpy(pts.ravel()))
            adj_h, det_h = S.adjugate_det(typ, 1, el, pts)
            np.testing.assert_array_equal(adj_d.numpy().reshape(-1, 9), adj_h)
            np.testing.assert_array_equal(det_d.numpy(), det_h)


@pytest.mark.parametrize("L", [2, 8])
def test_device_geometry_and_operators_match_host(L):
    mel, mpts = permuted_box(3, 7)
    el_h, pts_h = S.sshex8_refine(L, mel, mpts)
    mesh, pts_d = S.sshex8_refine_device(L, mel, mpts)
    N, E = pts_h.shape[1], mel.shape[1]
    fff_h = S.fff(S.SSHEX8, L, el_h, pts_h)
    fff_d = S.mesh_fff(mesh, pts_d)
    np.testing.assert_array_equal(fff_d.numpy().reshape(E, 6), fff_h)
    adj_h, det_h = S.adjugate_det(S.SSHEX8, L, el_h, pts_h)
    adj_d, det_d = S.mesh_adjugate_det(mesh, pts_d)
    np.testing.assert_array_equal(adj_d.numpy().reshape(E, 9), adj_h)
    np.testing.assert_array_equal(det_d.numpy(), det_h)
    # operators from device geometry vs the same words from the host
    x = S.DeviceArray.from_numpy(np.random.default_rng(L).uniform(-1, 1, N))
    lap_d = S.laplacian(mesh, fff_d)
    lap_h = S.laplacian(mesh, fff_d.numpy().reshape(E, 6))
    ya, yb = S.DeviceArray(N), S.DeviceArray(N)
    lap_d.apply(x, ya)
    lap_h.apply(x, yb)
    np.testing.assert_array_equal(ya.numpy(), yb.numpy())
    lap32_d = S.laplacian(mesh, fff_d, real_type=S.FLOAT32)
    lap32_h = S.laplacian(mesh, fff_d.numpy().reshape(E, 6), real_type=S.FLOAT32)
    x32 = S.DeviceArray.from_numpy(np.random.default_rng(L).uniform(-1, 1, N).astype(np.float32))
    za, zb = S.DeviceArray(N, np.float32), S.DeviceArray(N, np.float32)
    lap32_d.apply(x32, za)
    lap32_h.apply(x32, zb)
    np.testing.assert_array_equal(za.numpy(), zb.numpy())
    mu, lam = O.lame(1.0, 0.3)
    xe = S.DeviceArray.from_numpy(np.random.default_rng(L + 1).uniform(-1, 1, 3 * N))
    for rt, dt in ((S.FLOAT64, np.float64), (S.FLOAT32, np.float32)):
        el_d = S.linear_elasticity(mesh, adj_d, det_d, mu, lam, real_type=rt)
        el_hh = S.linear_elasticity(mesh, adj_d.numpy().reshape(E, 9), det_d.numpy(), mu, lam, real_type=rt)
        xv = xe if rt == S.FLOAT64 else S.DeviceArray.from_numpy(xe.numpy().astype(np.float32))
        ua, ub = S.DeviceArray(3 * N, dt), S.DeviceArray(3 * N, dt)
        el_d.apply(xv, ua)
        el_hh.apply(xv, ub)
        np.testing.assert_array_equal(ua.numpy(), ub.numpy())


def test_device_refine_setup_time_stated():
    """The setup the device path replaces: a 32^3 macro cube at L=8 (17.0M nodes)."""
    mel, mpts = S.box_mesh(33, 33, 33, element_type=S.HEX8)
    t0 = time.perf_counter()
    el_h, pts_h = S.sshex8_refine(8, mel, mpts)
    t_host = time.perf_counter() - t0
    me = S.DeviceArray.from_numpy(np.ascontiguousarray(mel, np.int32))
    mp = S.DeviceArray.from_numpy(np.ascontiguousarray(mpts, np.float32))
    S.sshex8_refine_device(8, me, mp)  # warm-up (code objects, allocator)
    S.synchronize()
    t0 = time.perf_counter()
    mesh, pts = S.sshex8_refine_device(8, me, mp)
    S.synchronize()
    t_dev = time.perf_counter() - t0
    print(f"sshex8 L=8 refinement of 32^3 macro elements ({mesh.nnodes} nodes): host {t_host:.3f} s, "
          f"device {t_dev:.3f} s")
    assert mesh.nnodes == pts_h.shape[1]
    np.testing.assert_array_equal(mesh.download()[:, ::997], el_h[:, ::997])
