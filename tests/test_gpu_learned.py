"""Learned smoother on the GPU: CNN forward and gradients against the float64 oracle
(oracle/learned_smoother.py), the fused BF16 MFMA inference against the FP32 path, training
from V-cycle pairs, the gain of the trained smoother in MG-PCG, and the Chebyshev3 fallback."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import sfem_amd as S  # noqa: E402
from sfem_amd import gmg, learned  # noqa: E402

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import learned_smoother as LS  # noqa: E402


def dev(a, dtype=np.float32):
    return S.DeviceArray.from_numpy(np.asarray(a), dtype)


def oracle_weights(seed, hidden=16, layers=4):
    """Oracle weights rounded to float32 (what the device holds) and their flat layout."""
    Ws = [w.astype(np.float32).astype(np.float64) for w in LS.init_weights(hidden, layers, seed)]
    return Ws, np.concatenate([w.ravel() for w in Ws]).astype(np.float32)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("L", [2, 4])
def test_cnn_forward_fp32_matches_oracle(L):
    n = (L + 1) ** 3
    E = 3
    Ws, flat = oracle_weights(1)
    z = np.random.default_rng(2).normal(size=(E, n)).astype(np.float32)
    ref, _ = LS.forward(z.astype(np.float64), Ws, L)
    cnn = S.CNN(16, 4, flat)
    out = S.DeviceArray(E * n, np.float32)
    cnn.forward(L, dev(z), out, mode=0)
    err = rel(out.numpy().reshape(E, n), ref)
    print(f"L={L} fp32 forward rel err {err:.2e}")
    assert err < 1e-5


def test_cnn_loss_and_gradient_match_oracle():
    L, E = 2, 3
    n = (L + 1) ** 3
    Ws, flat = oracle_weights(3)
    rng = np.random.default_rng(4)
    z = rng.normal(size=(E, n)).astype(np.float32)
    t = rng.normal(size=(E, n)).astype(np.float32)
    loss_ref, grads_ref = LS.loss_and_grads(z.astype(np.float64), t.astype(np.float64), Ws, L)
    cnn = S.CNN(16, 4, flat)
    loss, g = cnn.loss_grad(L, dev(z), dev(t))
    assert abs(loss - loss_ref) / loss_ref < 1e-5
    off = 0
    for i, gr in enumerate(grads_ref):
        gi = g[off:off + gr.size].reshape(gr.shape)
        off += gr.size
        err = rel(gi, gr)
        print(f"layer {i}: grad rel err {err:.2e}")
        assert err < 1e-4


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 6, 7, 8])
def test_cnn_mfma_inference_matches_fp32(L):
    n = (L + 1) ** 3
    E = 7
    Ws, flat = oracle_weights(5)
    z = np.random.default_rng(6).normal(size=(E, n)).astype(np.float32)
    cnn = S.CNN(16, 4, flat)
    o32 = S.DeviceArray(E * n, np.float32)
    o16 = S.DeviceArray(E * n, np.float32)
    cnn.forward(L, dev(z), o32, mode=0)
    cnn.forward(L, dev(z), o16, mode=1)
    a, b = o16.numpy(), o32.numpy()
    err = rel(a, b)
    print(f"L={L} MFMA BF16 vs FP32 rel err {err:.2e}")
    assert np.all(np.isfinite(a))
    assert err < 3e-2
    # oddness survives the BF16 path exactly: f(-z) = -f(z)
    o16m = S.DeviceArray(E * n, np.float32)
    cnn.forward(L, dev(-z), o16m, mode=1)
    np.testing.assert_array_equal(o16m.numpy(), -a)


def _poisson_problem(L=8, n_macro=4, coeff_seed=7):
    mel, mpts = S.box_mesh(n_macro + 1, n_macro + 1, n_macro + 1, element_type=S.HEX8)
    el, pts = S.sshex8_refine(L, mel, mpts)
    fff = S.fff(S.SSHEX8, L, el, pts)
    mesh = S.Mesh(S.SSHEX8, L, el, pts.shape[1])
    bc = np.where((pts.min(0) < 1e-6) | (pts.max(0) > 1 - 1e-6))[0].astype(np.int32)
    coeff = None
    if coeff_seed is not None:
        coeff = (10.0 ** np.random.default_rng(coeff_seed).uniform(0, 3, el.shape[1])).astype(np.float32)
    return mesh, fff, bc, coeff, pts.shape[1]


def _train(h, seed=0, steps=300):
    data = learned.collect_pairs(h, n_problems=4, cycles=2, seed=seed)
    cnn = S.CNN(16, 4, seed=seed)
    before = learned.evaluate(cnn, data)
    learned.train(cnn, data, steps=steps, lr=2e-3, batch_elements=16, seed=seed)
    after = learned.evaluate(cnn, data)
    return cnn, data, before, after


def test_training_from_vcycle_pairs_reduces_loss():
    mesh, fff, bc, coeff, N = _poisson_problem(L=8, n_macro=2, coeff_seed=None)
    h = gmg.SSGMG(mesh, "laplacian", fff, bc, coeff=coeff, real_type=S.FLOAT32)
    cnn, data, before, after = _train(h, steps=200)
    after_mfma = learned.evaluate(cnn, data, mode=1)
    print(f"pairs {len(data)}; loss before {before:.3f}, after {after:.3f} (MFMA inference {after_mfma:.3f})")
    assert after < 0.5 * before and after < 0.9
    assert abs(after_mfma - after) < 0.05


def _solve(h, N, rhs, A64, switches=(), ratio=0.5, rtol=1e-8):
    P = S.precision_bridge(h.mg) if h.real_type == S.FLOAT32 else h.mg
    solver = learned.flexible_cg_with_fallback(A64, P, list(switches), rtol=rtol, max_it=300, fallback_ratio=ratio)
    b = S.DeviceArray.from_numpy(rhs)
    x = S.DeviceArray(N)
    solver.apply(b, x)
    r = S.DeviceArray(N)
    A64.apply(x, r)
    res = float(np.linalg.norm(rhs - r.numpy()) / np.linalg.norm(rhs))
    return solver.iterations, solver.last_status, res


def test_learned_smoother_gain_over_chebyshev():
    """sshex8 Poisson (L=8, 4^3 macro elements, zero Dirichlet on all faces).  In the same FP32
    V-cycle inside FP64 flexible CG, two learned-smoother steps per smoothing (2 operator applies
    + 2 CNN evaluations) against Chebyshev3 of degree 3 (3 applies) on the finest level: fewer
    iterations to 1e-8 and no fallback.  The untrained network is caught by the fallback."""
    mesh, fff, bc, coeff, N = _poisson_problem(L=8, n_macro=4, coeff_seed=None)
    h_plain = gmg.SSGMG(mesh, "laplacian", fff, bc, real_type=S.FLOAT32)
    cnn, data, before, after = _train(h_plain, seed=1, steps=300)
    print(f"training loss {before:.3f} -> {after:.3f}")
    assert after < 0.1
    A64 = gmg.SSGMG(mesh, "laplacian", fff, bc).fine.A
    rhs = np.random.default_rng(8).uniform(0, 1, N)
    rhs[bc] = 0

    its_cheb, st, res = _solve(h_plain, N, rhs, A64)
    assert st == 0 and res < 1e-8
    h_learn = gmg.SSGMG(mesh, "laplacian", fff, bc, real_type=S.FLOAT32, learned=cnn, learned_steps=2)
    its_learn, st, res = _solve(h_learn, N, rhs, A64, h_learn.switches, ratio=0.5)
    _, events = S.switch_state(h_learn.switches[0])
    untrained = S.CNN(16, 4, seed=123)
    h_u = gmg.SSGMG(mesh, "laplacian", fff, bc, real_type=S.FLOAT32, learned=untrained, learned_steps=2)
    its_u, st_u, res_u = _solve(h_u, N, rhs, A64, h_u.switches, ratio=0.5)
    _, events_u = S.switch_state(h_u.switches[0])
    print(f"iterations: chebyshev3 {its_cheb}, learned x2 {its_learn} (fallbacks {events}), "
          f"untrained x2 {its_u} (fallbacks {events_u})")
    assert st == 0 and res < 1e-8 and events == 0
    assert its_learn < its_cheb
    assert events_u == 1 and st_u == 0 and res_u < 1e-8


def test_learned_smoother_variable_coefficient_falls_back():
    """Per-element coefficients 10^U[0,3]: Chebyshev3 on A stalls, Chebyshev3 on D^-1 A does
    not; a learned smoother with the scaled Chebyshev3 fallback converges like the latter."""
    mesh, fff, bc, coeff, N = _poisson_problem(L=8, n_macro=4, coeff_seed=7)
    h_cj = gmg.SSGMG(mesh, "laplacian", fff, bc, coeff=coeff, real_type=S.FLOAT32, smoother="cheb3_jacobi")
    cnn, data, before, after = _train(h_cj, seed=1, steps=300)
    A64 = gmg.SSGMG(mesh, "laplacian", fff, bc, coeff=coeff).fine.A
    rhs = np.random.default_rng(8).uniform(0, 1, N)
    rhs[bc] = 0
    its_cj, st, res = _solve(h_cj, N, rhs, A64)
    assert st == 0 and res < 1e-8
    h_l = gmg.SSGMG(mesh, "laplacian", fff, bc, coeff=coeff, real_type=S.FLOAT32, smoother="cheb3_jacobi",
                    learned=cnn, learned_steps=2)
    its_l, st_l, res_l = _solve(h_l, N, rhs, A64, h_l.switches, ratio=0.5)
    _, events = S.switch_state(h_l.switches[0])
    print(f"variable coefficient: scaled chebyshev3 {its_cj} its; learned x2 {its_l} its "
          f"(fallbacks {events}); training loss {before:.3f} -> {after:.3f}")
    assert st_l == 0 and res_l < 1e-8
    assert its_l <= its_cj + 3


def test_forced_fallback_switches_once():
    mesh, fff, bc, coeff, N = _poisson_problem(L=4, n_macro=2, coeff_seed=None)
    cnn = S.CNN(16, 4, seed=0)
    h = gmg.SSGMG(mesh, "laplacian", fff, bc, real_type=S.FLOAT32, learned=cnn)
    h64 = gmg.SSGMG(mesh, "laplacian", fff, bc)
    rhs = np.random.default_rng(9).uniform(0, 1, N)
    rhs[bc] = 0
    # ratio 0: the first iteration always switches
    its, st, res = _solve(h, N, rhs, h64.fine.A, h.switches, ratio=0.0)
    active, events = S.switch_state(h.switches[0])
    assert (active, events) == (1, 1)
    assert st == 0 and res < 1e-8
    # a switched hierarchy can be reset to the learned smoother
    h.select_smoothers(0)
    assert S.switch_state(h.switches[0])[0] == 0


def test_graph_replayed_learned_hierarchy_follows_switches_and_fallback():
    """A V-cycle replayed as a HIP graph must not replay stale host decisions: after the switch
    moves from the learned smoother to its Chebyshev3 fallback (by hand, or by the CG's fallback
    hook) the graph is re-captured, and its output equals the eager V-cycle in the same state.
    A flexible CG with fallback ratio 0 gives the same iterations and fallback events with the
    graph-replayed preconditioner as with the eager one."""
    mesh, fff, bc, coeff, N = _poisson_problem(L=4, n_macro=2, coeff_seed=None)
    cnn = S.CNN(16, 4, seed=0)
    h = gmg.SSGMG(mesh, "laplacian", fff, bc, real_type=S.FLOAT32, learned=cnn, coarse="dense")
    P_eager = h.preconditioner(graph=False)
    P_graph = h.preconditioner(graph=True)
    rng = np.random.default_rng(4)
    r = S.DeviceArray.from_numpy(rng.uniform(-1, 1, N))

    def both():
        ze, zg = S.DeviceArray(N), S.DeviceArray(N)
        P_eager.apply(r, ze)
        P_graph.apply(r, zg)
        S.synchronize()
        return ze.numpy(), zg.numpy()

    outs = {}
    for which in (0, 1, 0):
        h.select_smoothers(which)
        for _ in range(3):  # first call eager, then captured, then replayed
            ze, zg = both()
            err = np.linalg.norm(zg - ze) / np.linalg.norm(ze)
            assert err < 1e-5, (which, err)
        outs.setdefault(which, ze)
    # the two smoothers give different cycles: a stale replay would be caught above
    assert np.linalg.norm(outs[0] - outs[1]) > 1e-3 * np.linalg.norm(outs[1])
    h64 = gmg.SSGMG(mesh, "laplacian", fff, bc)
    rhs = np.random.default_rng(9).uniform(0, 1, N)
    rhs[bc] = 0
    results = []
    for P in (P_eager, P_graph):
        h.select_smoothers(0)
        solver = learned.flexible_cg_with_fallback(h64.fine.A, P, h.switches, rtol=1e-8, max_it=300,
                                                   fallback_ratio=0.0)
        x = S.DeviceArray(N)
        solver.apply(S.DeviceArray.from_numpy(rhs), x)
        active, events = S.switch_state(h.switches[0])
        results.append((solver.iterations, solver.last_status, active))
    print(f"eager vs graph-replayed (iterations, status, active smoother): {results}")
    assert results[0] == results[1]
    assert results[0][1] == 0 and results[0][2] == 1


def test_vector_cnn_forward_and_gradient_match_oracle():
    """The 3-channel network (one channel per displacement component) against the float64 oracle:
    forward and the loss gradient of every layer."""
    L, E = 2, 3
    n = (L + 1) ** 3
    Ws = [w.astype(np.float32).astype(np.float64) for w in LS.init_weights(16, 4, 11, channels=3)]
    flat = np.concatenate([w.ravel() for w in Ws]).astype(np.float32)
    rng = np.random.default_rng(12)
    z = rng.normal(size=(E, 3, n)).astype(np.float32)
    t = rng.normal(size=(E, 3, n)).astype(np.float32)
    cnn = S.CNN(16, 4, flat, channels=3)
    assert cnn.num_params == LS.n_params(16, 4, channels=3)
    ref, _ = LS.forward(z.astype(np.float64), Ws, L)
    out = S.DeviceArray(E * 3 * n, np.float32)
    cnn.forward(L, dev(z), out, mode=0)
    err = rel(out.numpy().reshape(E, 3, n), ref)
    loss_ref, grads_ref = LS.loss_and_grads(z.astype(np.float64), t.astype(np.float64), Ws, L)
    loss, g = cnn.loss_grad(L, dev(z), dev(t))
    print(f"vector network: forward rel err {err:.2e}, loss {loss:.6f} vs {loss_ref:.6f}")
    assert err < 1e-5
    assert abs(loss - loss_ref) / loss_ref < 1e-5
    off = 0
    for i, gr in enumerate(grads_ref):
        gi = g[off:off + gr.size].reshape(gr.shape)
        off += gr.size
        assert rel(gi, gr) < 1e-4, i


def _elasticity_problem(L=8, n_macro=2):
    mel, mpts = S.box_mesh(n_macro + 1, n_macro + 1, n_macro + 1, element_type=S.HEX8)
    el, pts = S.sshex8_refine(L, mel, mpts)
    adj, det = S.adjugate_det(S.HEX8, 1, mel, mpts)
    mesh = S.Mesh(S.SSHEX8, L, el, pts.shape[1])
    bc = np.where(pts[2] < 1e-6)[0].astype(np.int32)
    nu = 0.3
    mu, lam = 1.0 / (2 * (1 + nu)), nu / ((1 + nu) * (1 - 2 * nu))
    N = pts.shape[1]
    rhs = np.zeros((N, 3))
    rhs[:, 2] = -1
    rhs[bc] = 0
    return mesh, (adj, det), bc, mu, lam, N, rhs.ravel()


def test_vector_learned_smoother_on_elasticity():
    """sshex8 linear elasticity (L=8, 2^3 macro elements, clamped at z = 0, body force): the 3-channel
    network trained on the plain V-cycle's (D^-1 r, e) pairs replaces Chebyshev3 on the finest level of
    the FP32 hierarchy inside FP64 flexible CG.  The loss drops by more than half, but the trained
    network does not yet smooth elasticity as well as Chebyshev3 (relative training loss ~0.1-0.17
    against ~0.02 for Poisson): the residual monitor falls back to Chebyshev3 once, and the solve then
    converges within two iterations of plain Chebyshev3.  The iteration counts of Chebyshev3, the
    vector network and the scalar network applied per component are reported."""
    mesh, geom, bc, mu, lam, N, rhs = _elasticity_problem()
    h32 = gmg.SSGMG(mesh, "elasticity", geom, bc, mu=mu, lam=lam, real_type=S.FLOAT32)
    A64 = gmg.SSGMG(mesh, "elasticity", geom, bc, mu=mu, lam=lam).fine.A
    data = learned.collect_pairs(h32, n_problems=4, cycles=2, seed=3, channels=3)
    assert data.channels == 3 and len(data) == 12
    cnn = S.CNN(16, 4, seed=3, channels=3)
    before = learned.evaluate(cnn, data)
    learned.train(cnn, data, steps=300, lr=2e-3, batch_elements=8, seed=3)
    after = learned.evaluate(cnn, data)
    its_cheb, st, res = _solve(h32, 3 * N, rhs, A64)
    assert st == 0 and res < 1e-8
    h_v = gmg.SSGMG(mesh, "elasticity", geom, bc, mu=mu, lam=lam, real_type=S.FLOAT32, learned=cnn, learned_steps=2)
    its_v, st_v, res_v = _solve(h_v, 3 * N, rhs, A64, h_v.switches, ratio=0.5)
    _, events = S.switch_state(h_v.switches[0])
    data1 = learned.collect_pairs(h32, n_problems=4, cycles=2, seed=3)
    cnn1 = S.CNN(16, 4, seed=3)
    learned.train(cnn1, data1, steps=300, lr=2e-3, batch_elements=8, seed=3)
    h_s = gmg.SSGMG(mesh, "elasticity", geom, bc, mu=mu, lam=lam, real_type=S.FLOAT32, learned=cnn1, learned_steps=2)
    its_s, st_s, res_s = _solve(h_s, 3 * N, rhs, A64, h_s.switches, ratio=0.5)
    _, events_s = S.switch_state(h_s.switches[0])
    print(f"elasticity: loss {before:.3f} -> {after:.3f}; iterations chebyshev3 {its_cheb}, vector network x2 "
          f"{its_v} (fallbacks {events}), scalar network per component x2 {its_s} (fallbacks {events_s})")
    assert after < 0.5 * before
    assert st_v == 0 and res_v < 1e-8 and events <= 1 and its_v <= its_cheb + 2
    assert st_s == 0 and res_s < 1e-8


def test_variable_coefficient_protocol_held_out_seeds():
    """The variable-coefficient protocol at reduced size (sshex8 Poisson, L=8, 4^3 macro elements):
    training pairs after one plain V-cycle on 24 seeded problems (log-normal piecewise-constant
    coefficients, log10 contrast U[0,3]), then on three held-out seeds the trained network (2 steps)
    against the scaled Chebyshev3 it falls back to: held-out loss below the untrained network's, every
    solve converged, no fallback, and the iteration counts reported."""
    L, n_macro = 8, 4
    mesh, fff, bc, _, N = _poisson_problem(L=L, n_macro=n_macro, coeff_seed=None)
    E = n_macro ** 3

    def build(seed):
        k, _ = learned.coefficient_field(seed, E)
        return gmg.SSGMG(mesh, "laplacian", fff, bc, coeff=k, real_type=S.FLOAT32, smoother="cheb3_jacobi")

    train_set = learned.protocol_pairs(build, range(24), cycles=1)
    test_set = learned.protocol_pairs(build, range(1000, 1003), cycles=1)
    cnn = S.CNN(16, 4, seed=5)
    untrained = learned.evaluate(cnn, test_set)
    learned.train(cnn, train_set, steps=400, lr=2e-3, batch_elements=16, seed=5)
    held_out = learned.evaluate(cnn, test_set)
    print(f"protocol: {len(train_set)} training pairs; held-out loss untrained {untrained:.3f} -> trained "
          f"{held_out:.3f}")
    assert held_out < untrained
    rows = []
    for seed in range(1000, 1003):
        k, c = learned.coefficient_field(seed, E)
        A64 = gmg.SSGMG(mesh, "laplacian", fff, bc, coeff=k).fine.A
        rhs = np.random.default_rng(seed).uniform(0, 1, N)
        rhs[bc] = 0
        h_cj = build(seed)
        its_cj, st, res = _solve(h_cj, N, rhs, A64)
        h_l = gmg.SSGMG(mesh, "laplacian", fff, bc, coeff=k, real_type=S.FLOAT32, smoother="cheb3_jacobi",
                        learned=cnn, learned_steps=2)
        its_l, st_l, res_l = _solve(h_l, N, rhs, A64, h_l.switches, ratio=0.5)
        _, events = S.switch_state(h_l.switches[0])
        rows.append((seed, round(c, 2), its_cj, its_l, events))
        assert st == 0 and res < 1e-8 and st_l == 0 and res_l < 1e-8
    print("held-out (seed, log10 contrast, scaled chebyshev3 its, learned x2 its, fallbacks):", rows)
    assert all(r[4] == 0 for r in rows)
