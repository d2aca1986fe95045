"""CPU checks of the product library: it loads, exports every symbol include/sfem_amd.h
declares, and its host-side mesh/geometry setup agrees with the oracle (connectivity and
numbering bit-exact, float32 geometry words within a few ulp).  No GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import sfem_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sfem_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sfem_amd_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    syms = declared_symbols()
    assert len(syms) > 40
    lib = ctypes.CDLL(S.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding covers exactly the declared ABI
    assert sorted(S.SIGNATURES) == syms


@pytest.mark.parametrize("tet4", [False, True])
def test_box_mesh_matches_oracle(tet4):
    for dims in ((3, 4, 5), (9, 9, 9)):
        a = S.box_mesh(*dims, 1.0, 2.0, 0.5, S.TET4 if tet4 else S.HEX8)
        b = O.box_mesh(*dims, 1.0, 2.0, 0.5, tet4)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("hierarchical", [True, False])
@pytest.mark.parametrize("n,L", [(2, 2), (3, 4), (2, 8), (2, 6), (1, 16), (2, 12), (2, 3)])
def test_sshex8_refine_matches_oracle_bitexact(n, L, hierarchical):
    """Product host refinement (sort-based edge/side ranks, first-element table) against the oracle's
    line-by-line restatement of sshex8_generate_elements + sshex8_hierarchical_renumbering, which the
    hand-derived tables of test_oracle_mesh.py pin."""
    mel, mpts = O.box_mesh(n + 1, n + 1, n + 1)
    e1, p1 = S.sshex8_refine(L, mel, mpts, hierarchical=hierarchical)
    e2, p2 = O.sshex8_refine(L, mel, mpts, hierarchical=hierarchical)
    np.testing.assert_array_equal(e1, e2)
    np.testing.assert_allclose(p1, p2, rtol=0, atol=1e-6)


@pytest.mark.parametrize("hierarchical", [True, False])
def test_sshex8_refine_permuted_elements_matches_oracle(hierarchical):
    mel, mpts = O.box_mesh(3, 4, 3)
    rng = np.random.default_rng(5)
    mel = mel[:, rng.permutation(mel.shape[1])]
    perm = [1, 2, 3, 0, 5, 6, 7, 4]
    mel[:, ::3] = mel[perm][:, ::3]
    for L in (3, 4):
        e1, _ = S.sshex8_refine(L, mel, mpts, hierarchical=hierarchical)
        e2, _ = O.sshex8_refine(L, mel, mpts, hierarchical=hierarchical)
        np.testing.assert_array_equal(e1, e2)


def _ulp_close(a, b, ulps=4):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    tol = ulps * np.spacing(np.maximum(np.abs(a), np.abs(b)))
    assert np.all(np.abs(a - b) <= tol + 1e-30), np.max(np.abs(a - b) / (np.abs(b) + 1e-30))


def test_geometry_matches_oracle():
    rng = np.random.default_rng(3)
    mel, mpts = O.box_mesh(4, 4, 4)
    mpts = (mpts + 0.05 * rng.uniform(-1, 1, mpts.shape)).astype(np.float32)  # non-affine
    _ulp_close(S.fff(S.HEX8, 1, mel, mpts), O.hex8_fff(mel, mpts))
    a1, d1 = S.adjugate_det(S.HEX8, 1, mel, mpts)
    a2, d2 = O.hex8_adj_det(mel, mpts)
    np.testing.assert_allclose(a1, a2, rtol=1e-6, atol=1e-7)
    _ulp_close(d1, d2)
    tel, tpts = O.box_mesh(4, 4, 4, tet4=True)
    tpts = (tpts + 0.05 * rng.uniform(-1, 1, tpts.shape)).astype(np.float32)
    # tet4: both follow the reference's float32 evaluation order: identical words
    np.testing.assert_array_equal(S.fff(S.TET4, 1, tel, tpts), O.tet4_fff(tel, tpts))
    # sshex8 macro geometry reads the macro corners of the refined connectivity
    L = 4
    el, pts = O.sshex8_refine(L, mel, mpts)
    _ulp_close(S.fff(S.SSHEX8, L, el, pts), O.hex8_fff(mel, mpts))


def test_errors_are_reported_not_fatal():
    with pytest.raises(S.SfemError):
        S.box_mesh(1, 3, 3)
