"""Pins the oracle's mesh generators.

Box meshes: bit-exact against the reference's own python/sfem/mesh/box_mesh.py output
(tests/golden/*.npz, written by tests/golden/make_box_mesh_golden.py).
sshex8 refinement: known-answer tables worked out by hand from the reference's numbering text
(mesh/sfem_hex8_mesh_graph.cpp:448-735 sshex8_generate_elements, :994-1060
sshex8_hierarchical_renumbering; the live smesh::to_semistructured is absent), plus invariants.
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


@pytest.mark.parametrize("n,L,levels", [(2, 8, [8, 4, 2, 1]), (2, 4, [4, 2, 1]), (2, 6, [6, 1]), (1, 16, [16, 8, 4, 2, 1]),
                                        (1, 12, [12, 6, 1])])
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


# ------------------------------------------------------------------ hand-derived numbering tables
def _table(el_col, L):
    """connectivity column -> nested [z][y][x] list"""
    return el_col.reshape(L + 1, L + 1, L + 1).tolist()


def test_hierarchical_levels_match_reference_text():
    """sshex8_hierarchical_mesh_levels (mesh/sfem_hex8_mesh_graph.cpp:964-992); [1, 2, 4, 8] is also
    the derefinement test of frontend/tests/sfem_SemiStructuredMeshTest.cpp:18-22.  Odd L > 1:
    the reference list is [1] and its renumbering aborts; the restatement appends L."""
    assert O.hierarchical_levels(8) == [1, 2, 4, 8]
    assert O.hierarchical_levels(16) == [1, 2, 4, 8, 16]
    assert O.hierarchical_levels(12) == [1, 6, 12]
    assert O.hierarchical_levels(6) == [1, 6]
    assert O.hierarchical_levels(1) == [1]
    assert O.hierarchical_levels(3) == [1, 3]


def test_numbering_kat_one_element_L2_hierarchical():
    """Corners keep 0..7; then the level-2 nodes in (z, y, x) first-encounter order take 8, 9, ...
    (the z=0 face centre is the third new node: 10)."""
    mel = np.arange(8, dtype=np.int32).reshape(8, 1)
    el, _ = O.sshex8_refine(2, mel, np.zeros((3, 8), np.float32))
    assert _table(el[:, 0], 2) == [[[0, 8, 1], [9, 10, 11], [3, 12, 2]],
                                   [[13, 14, 15], [16, 17, 18], [19, 20, 21]],
                                   [[4, 22, 5], [23, 24, 25], [7, 26, 6]]]


def test_numbering_kat_one_element_L2_raw():
    """sshex8_generate_elements without renumbering: edges in CRS (i<j) order from 8, sides in
    side-table order (python/codegen/hex8.py:44-52: y=0, x=L, y=L, x=0, z=0, z=L) from 20, the
    volume node 26."""
    mel = np.arange(8, dtype=np.int32).reshape(8, 1)
    el, _ = O.sshex8_refine(2, mel, np.zeros((3, 8), np.float32), hierarchical=False)
    assert _table(el[:, 0], 2) == [[[0, 8, 1], [9, 24, 11], [3, 13, 2]],
                                   [[10, 20, 12], [23, 26, 21], [15, 22, 14]],
                                   [[4, 16, 5], [17, 25, 18], [7, 19, 6]]]


# L = 3, one element with permuted corner ids: edges run from the smaller id (reversed walks), the
# side origin is the smallest id and u its smaller neighbour (swaps).  (x, y, z) -> id.
_KAT_L3_CORNERS = [5, 2, 7, 0, 6, 1, 3, 4]
_KAT_L3_RAW = {
    # corners
    (0, 0, 0): 5, (3, 0, 0): 2, (3, 3, 0): 7, (0, 3, 0): 0, (0, 0, 3): 6, (3, 0, 3): 1, (3, 3, 3): 3, (0, 3, 3): 4,
    # edges (ids 8 + 2*edge + step from the smaller corner id)
    (0, 3, 1): 8, (0, 3, 2): 9,        # (0,4)
    (0, 2, 0): 10, (0, 1, 0): 11,      # (0,5)
    (1, 3, 0): 12, (2, 3, 0): 13,      # (0,7)
    (3, 0, 2): 14, (3, 0, 1): 15,      # (1,2)
    (3, 1, 3): 16, (3, 2, 3): 17,      # (1,3)
    (2, 0, 3): 18, (1, 0, 3): 19,      # (1,6)
    (2, 0, 0): 20, (1, 0, 0): 21,      # (2,5)
    (3, 1, 0): 22, (3, 2, 0): 23,      # (2,7)
    (2, 3, 3): 24, (1, 3, 3): 25,      # (3,4)
    (3, 3, 2): 26, (3, 3, 1): 27,      # (3,7)
    (0, 2, 3): 28, (0, 1, 3): 29,      # (4,6)
    (0, 0, 1): 30, (0, 0, 2): 31,      # (5,6)
    # sides from 32, four nodes each, u fastest
    (2, 0, 2): 32, (2, 0, 1): 33, (1, 0, 2): 34, (1, 0, 1): 35,   # y=0: o=1 u=2 v=6
    (3, 1, 2): 36, (3, 1, 1): 37, (3, 2, 2): 38, (3, 2, 1): 39,   # x=3: o=1 u=2 v=3
    (1, 3, 1): 40, (1, 3, 2): 41, (2, 3, 1): 42, (2, 3, 2): 43,   # y=3: o=0 u=4 v=7
    (0, 2, 1): 44, (0, 2, 2): 45, (0, 1, 1): 46, (0, 1, 2): 47,   # x=0: o=0 u=4 v=5
    (1, 2, 0): 48, (1, 1, 0): 49, (2, 2, 0): 50, (2, 1, 0): 51,   # z=0: o=0 u=5 v=7
    (2, 1, 3): 52, (2, 2, 3): 53, (1, 1, 3): 54, (1, 2, 3): 55,   # z=3: o=1 u=3 v=6
    # volume from 56, x fastest
    (1, 1, 1): 56, (2, 1, 1): 57, (1, 2, 1): 58, (2, 2, 1): 59,
    (1, 1, 2): 60, (2, 1, 2): 61, (1, 2, 2): 62, (2, 2, 2): 63,
}


def test_numbering_kat_permuted_element_L3_raw():
    mel = np.array(_KAT_L3_CORNERS, dtype=np.int32).reshape(8, 1)
    el, _ = O.sshex8_refine(3, mel, np.zeros((3, 8), np.float32), hierarchical=False)
    assert len(_KAT_L3_RAW) == 64
    for (x, y, z), want in _KAT_L3_RAW.items():
        assert el[O.lidx(3, x, y, z), 0] == want, (x, y, z)


def test_numbering_kat_two_elements_L2_hierarchical():
    """Two elements side by side (grid ids x + 3y + 6z): element 1 reuses the shared side's ids
    given by element 0 and numbers its own nodes after element 0's."""
    mel = np.array([[0, 1, 4, 3, 6, 7, 10, 9], [1, 2, 5, 4, 7, 8, 11, 10]], dtype=np.int32).T.copy()
    el, _ = O.sshex8_refine(2, mel, np.zeros((3, 12), np.float32))
    assert _table(el[:, 0], 2) == [[[0, 12, 1], [13, 14, 15], [3, 16, 4]],
                                   [[17, 18, 19], [20, 21, 22], [23, 24, 25]],
                                   [[6, 26, 7], [27, 28, 29], [9, 30, 10]]]
    assert _table(el[:, 1], 2) == [[[1, 31, 2], [15, 32, 33], [4, 34, 5]],
                                   [[19, 35, 36], [22, 37, 38], [25, 39, 40]],
                                   [[7, 41, 8], [29, 42, 43], [10, 44, 11]]]


def test_numbering_kat_one_element_L4_hierarchical():
    """Levels [1, 2, 4]: the 19 new level-2 nodes take 8..26 in (z, y, x) order, then the level-4
    nodes from 27; the last new node, (3, 4, 4), takes 124."""
    mel = np.arange(8, dtype=np.int32).reshape(8, 1)
    el, _ = O.sshex8_refine(4, mel, np.zeros((3, 8), np.float32))
    got = lambda x, y, z: int(el[O.lidx(4, x, y, z), 0])
    level2 = [(2, 0, 0), (0, 2, 0), (2, 2, 0), (4, 2, 0), (2, 4, 0),
              (0, 0, 2), (2, 0, 2), (4, 0, 2), (0, 2, 2), (2, 2, 2), (4, 2, 2), (0, 4, 2), (2, 4, 2), (4, 4, 2),
              (2, 0, 4), (0, 2, 4), (2, 2, 4), (4, 2, 4), (2, 4, 4)]
    assert [got(*p) for p in level2] == list(range(8, 27))
    plane0 = [(1, 0, 0), (3, 0, 0), (0, 1, 0), (1, 1, 0), (2, 1, 0), (3, 1, 0), (4, 1, 0), (1, 2, 0), (3, 2, 0),
              (0, 3, 0), (1, 3, 0), (2, 3, 0), (3, 3, 0), (4, 3, 0), (1, 4, 0), (3, 4, 0)]
    assert [got(*p) for p in plane0] == list(range(27, 43))
    assert got(0, 0, 1) == 43 and got(3, 4, 4) == 124 and el.max() == 124
