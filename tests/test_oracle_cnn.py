"""Pins the numpy restatement of the learned smoother (oracle/learned_smoother.py)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import learned_smoother as LS  # noqa: E402


def brute_conv(x, W, L):
    n1 = L + 1
    E, Ci, n = x.shape
    out = np.zeros((E, W.shape[0], n))
    g = x.reshape(E, Ci, n1, n1, n1)
    for z in range(n1):
        for y in range(n1):
            for xx in range(n1):
                for k, (dz, dy, dx) in enumerate(LS.OFFSETS):
                    zz, yy, xs = z + dz, y + dy, xx + dx
                    if 0 <= zz < n1 and 0 <= yy < n1 and 0 <= xs < n1:
                        out[:, :, (z * n1 + y) * n1 + xx] += g[:, :, zz, yy, xs] @ W[:, :, k].T
    return out


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    L = 3
    x = rng.normal(size=(2, 3, (L + 1) ** 3))
    W = rng.normal(size=(4, 3, 27))
    np.testing.assert_allclose(LS.conv(x, W, L), brute_conv(x, W, L), rtol=1e-12, atol=1e-12)


def test_odd_and_homogeneous():
    rng = np.random.default_rng(1)
    L = 4
    Ws = LS.init_weights(8, 4, seed=3)
    z = rng.normal(size=(3, (L + 1) ** 3))
    f, _ = LS.forward(z, Ws, L)
    np.testing.assert_allclose(LS.forward(-z, Ws, L)[0], -f, atol=1e-12)
    np.testing.assert_allclose(LS.forward(2.5 * z, Ws, L)[0], 2.5 * f, rtol=1e-12, atol=1e-12)
    assert LS.n_params(16, 4) == 27 * (16 + 256 + 256 + 16)


def test_gradients_by_finite_differences():
    rng = np.random.default_rng(2)
    L = 2
    Ws = LS.init_weights(4, 4, seed=5)
    z = rng.normal(size=(2, (L + 1) ** 3))
    t = rng.normal(size=(2, (L + 1) ** 3))
    loss, grads = LS.loss_and_grads(z, t, Ws, L)
    eps = 1e-6
    for li in range(len(Ws)):
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in Ws[li].shape)
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[li][idx] += eps
            Wm[li][idx] -= eps
            fd = (LS.loss_and_grads(z, t, Wp, L)[0] - LS.loss_and_grads(z, t, Wm, L)[0]) / (2 * eps)
            assert abs(fd - grads[li][idx]) <= 1e-6 * max(1.0, abs(fd)), (li, idx, fd, grads[li][idx])


def test_vector_network_odd_homogeneous_and_gradients():
    """The 3-channel network (one channel per displacement component): odd, positively homogeneous,
    gradients by finite differences, and with block-diagonal weights (no channel coupling) it is three
    independent scalar networks."""
    rng = np.random.default_rng(6)
    L = 2
    n = (L + 1) ** 3
    Ws = LS.init_weights(4, 3, seed=7, channels=3)
    assert [w.shape for w in Ws] == [(4, 3, 27), (4, 4, 27), (3, 4, 27)]
    assert LS.n_params(4, 3, channels=3) == 27 * (12 + 16 + 12)
    z = rng.normal(size=(2, 3, n))
    t = rng.normal(size=(2, 3, n))
    f, _ = LS.forward(z, Ws, L)
    assert f.shape == (2, 3, n)
    np.testing.assert_allclose(LS.forward(-z, Ws, L)[0], -f, atol=1e-12)
    np.testing.assert_allclose(LS.forward(0.7 * z, Ws, L)[0], 0.7 * f, rtol=1e-12, atol=1e-12)
    loss, grads = LS.loss_and_grads(z, t, Ws, L)
    eps = 1e-6
    for li in range(len(Ws)):
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in Ws[li].shape)
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[li][idx] += eps
            Wm[li][idx] -= eps
            fd = (LS.loss_and_grads(z, t, Wp, L)[0] - LS.loss_and_grads(z, t, Wm, L)[0]) / (2 * eps)
            assert abs(fd - grads[li][idx]) <= 1e-6 * max(1.0, abs(fd)), (li, idx, fd, grads[li][idx])
    # channel-diagonal weights: hidden width 3, channel c only talks to channel c
    Wd = []
    for W in LS.init_weights(3, 3, seed=8, channels=3):
        W = W.copy()
        for a in range(3):
            for b in range(3):
                if a != b:
                    W[a, b] = 0
        Wd.append(W)
    fv, _ = LS.forward(z, Wd, L)
    for c in range(3):
        fs, _ = LS.forward(z[:, c, :], [W[c:c + 1, c:c + 1] for W in Wd], L)
        np.testing.assert_allclose(fv[:, c, :], fs, rtol=1e-12, atol=1e-12)
