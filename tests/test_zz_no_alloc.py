"""The allocation-free dot (sfem_amd_blas_dot_noalloc) against the allocating one (sfem_amd_blas_dot).

Both run the same kernels on the same fixed reduction grid and sum the partials in the same
order, so the results must be the same word (integer-style bar: bit-exact).  Two calls in a row,
the second with a smaller input, check that the shared static scratch array carries nothing from
one call into the next; an empty vector gives 0.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import sfem_amd as S  # noqa: E402


def _dot_noalloc(x, y, result):
    rt = S.FLOAT64 if x.element_size() == 8 else S.FLOAT32
    rc = S.lib.sfem_amd_blas_dot_noalloc(rt, x.numel(), C.c_void_p(x.ptr), C.c_void_p(y.ptr),
                                         C.c_void_p(result.ptr), None)
    assert rc == S.SUCCESS, S.lib.sfem_amd_last_error().decode()
    S.synchronize()
    return float(result.numpy()[0])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dot_noalloc_matches_dot_on_two_calls(dtype):
    rng = np.random.default_rng(7)
    result = S.DeviceArray(1, np.float64)
    # first call large (every one of the 1024 partial slots in use), second call smaller
    for n in (3_000_001, 12_345, 0):
        xh = rng.uniform(-1, 1, n).astype(dtype)
        yh = rng.uniform(-1, 1, n).astype(dtype)
        x, y = S.DeviceArray.from_numpy(xh), S.DeviceArray.from_numpy(yh)
        ref = S.lib.sfem_amd_blas_dot(S.FLOAT64 if dtype == np.float64 else S.FLOAT32, n,
                                      C.c_void_p(x.ptr), C.c_void_p(y.ptr), None)
        got = _dot_noalloc(x, y, result)
        print(f"dot_noalloc dtype={np.dtype(dtype).name} n={n}: {got!r} vs {ref!r}", flush=True)
        assert got == ref  # same word
        exact = float(np.dot(xh.astype(np.float64), yh.astype(np.float64)))
        assert abs(got - exact) <= 1e-12 * max(1.0, float(np.abs(xh.astype(np.float64)) @ np.abs(yh.astype(np.float64))))


def test_dot_noalloc_rejects_bad_arguments():
    result = S.DeviceArray(1, np.float64)
    x = S.DeviceArray.from_numpy(np.ones(4))
    rc = S.lib.sfem_amd_blas_dot_noalloc(S.FLOAT64, -1, C.c_void_p(x.ptr), C.c_void_p(x.ptr),
                                         C.c_void_p(result.ptr), None)
    assert rc == S.FAILURE
    assert "n must be >= 0" in S.lib.sfem_amd_last_error().decode()
    rc = S.lib.sfem_amd_blas_dot_noalloc(S.FLOAT64, 4, C.c_void_p(x.ptr), C.c_void_p(x.ptr), None, None)
    assert rc == S.FAILURE
