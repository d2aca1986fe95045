"""Pins the oracle's mesh generators.

Box meshes: bit-exact against the reference's own python/sfem/mesh/box_mesh.py output
(tests/golden/*.npz, written by tests/golden/make_box_mesh_golden.py).
sshex8 refinement: invariants (the live numbering is in the absent smesh submodule).
"""
import glob
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "box_*.npz"))))
def test_box_mesh_matches_reference_generator(path):
    g = np.load(path)
    nx, ny, nz = (int(v) for v in g["dims"])
    w, h, t = (float(v) for v in g["extent"])
    tet4 = "tet4" in os.path.basename(path)
    el, pts = O.box_mesh(nx, ny, nz, w, h, t, tet4)
    np.testing.assert_array_equal(el, g["elements"])
    np.testing.assert_array_equal(pts, g["points"])


def _structured_ss(n, L):
    mel, mpts = O.box_mesh(n + 1, n + 1, n + 1)
    el, pts = O.sshex8_refine(L, mel, mpts)
    return mel, mpts, el, pts


@pytest.mark.parametrize("n,L", [(1, 2), (2, 4), (3, 4), (2, 8), (2, 6)])
def test_sshex8_refine_counts_and_coordinates(n, L):
    mel, mpts, el, pts = _structured_ss(n, L)
    N = pts.shape[1]
    assert N == (n * L + 1) ** 3
    # every id used, each exactly where the structured grid puts it
    assert el.min() == 0 and el.max() == N - 1
    assert len(np.unique(el)) == N
    # corners keep the macro ids
    np.testing.assert_array_equal(el[O.sshex8_corner_rows(L)], mel[:, :])
    # coordinates of each local node equal the trilinear image of the macro corners
    h = 1.0 / L
    for e in range(el.shape[1]):
        c = mpts[:, mel[:, e]].astype(np.float64)
        for z in range(L + 1):
            for y in range(L + 1):
                for x in range(L + 1):
                    q = (x * h, y * h, z * h)
                    f = np.array([(q[0] if cx else 1 - q[0]) * (q[1] if cy else 1 - q[1]) * (q[2] if cz else 1 - q[2])
                                  for (cx, cy, cz) in O.HEX8_CORNERS])
                    np.testing.assert_allclose(pts[:, el[O.lidx(L, x, y, z), e]], c @ f, atol=1e-6)


@pytest.mark.parametrize("n,L,levels", [(2, 8, [8, 4, 2, 1]), (2, 4, [4, 2, 1]), (2, 6, [6, 3, 1]), (1, 16, [16, 8, 4, 2, 1])])
def test_sshex8_hierarchical_prefix_property(n, L, levels):
    """Every derefined level is a prefix of the fine numbering
    (frontend/sfem_FunctionSpace.cpp:178-188 sizes a level as max id + 1;
    frontend/tests/sfem_SemiStructuredMeshTest.cpp:18-22 levels [1,2,4,8])."""
    mel, mpts, el, pts = _structured_ss(n, L)
    for l in levels[1:]:
        c = O.derefine(L, l, el)
        nl = (n * l + 1) ** 3
        assert c.max() + 1 == nl
        assert len(np.unique(c)) == nl
    h8 = O.to_hex8(1, O.derefine(L, 1, el))
    np.testing.assert_array_equal(h8, mel)


def test_sshex8_refine_unstructured_orientation():
    """Refinement of a mesh whose elements are rotated/reflected copies still
    conforms: shared nodes get one id with one coordinate."""
    mel, mpts = O.box_mesh(3, 3, 2)
    # rotate every other element's local numbering by a face-preserving symmetry
    perm = [1, 2, 3, 0, 5, 6, 7, 4]
    mel2 = mel.copy()
    mel2[:, 1::2] = mel[perm][:, 1::2]
    L = 4
    el, pts = O.sshex8_refine(L, mel2, mpts)
    N = pts.shape[1]
    assert N == (2 * L + 1) * (2 * L + 1) * (1 * L + 1)
    # coordinates unique per id: no two ids at the same point
    key = np.round(pts.T * 1e5).astype(np.int64)
    assert len(np.unique(key, axis=0)) == N


def test_incidence_counts():
    mel, mpts, el, pts = _structured_ss(2, 4)
    c = O.incidence(el, pts.shape[1])
    assert c.sum() == el.size
    # domain corner nodes belong to 1 macro element, the centre of the 2^3 block to 8
    corner = np.where((pts[0] == 0) & (pts[1] == 0) & (pts[2] == 0))[0][0]
    centre = np.where((np.abs(pts[0] - 0.5) < 1e-6) & (np.abs(pts[1] - 0.5) < 1e-6) & (np.abs(pts[2] - 0.5) < 1e-6))[0][0]
    assert c[corner] == 1 and c[centre] == 8
