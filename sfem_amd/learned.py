"""Learned smoother for sshex8 multigrid: training data from plain V-cycles, training on the
device, and hierarchies whose fine levels smooth with the CNN.

No counterpart in the reference: this is chosen form 4.  The smoother on an sshex8 level with
micro-grid level L <= 8 is

    x += damping * avg_e f(D^-1 (b - A x)|_e),     f(z) = g(z) - g(-z)

with g a 4-layer 3x3x3 CNN (16 channels, no biases) applied to each macro element's (L+1)^3
micro-grid.  f is odd and positively homogeneous of degree 1, so it commutes with scaling
of the residual.

Training pairs come from plain V-cycles (Chebyshev3 smoothing) on the homogeneous problem
A x = 0 started from random initial guesses, where the error is e = -x: after j = 0..cycles
V-cycles, z = D^-1 r with r = -A x_j and the target is the element restriction of e.  Pairs
after 0 cycles hold rough error (what a smoother must remove); later pairs hold the error the
V-cycle leaves behind.  The loss per minibatch (a range of macro elements of one pair) is
sum (f(z) - e)^2 / sum e^2, minimised with Adam.

The fallback: in a flexible CG, an iteration whose residual reduction ||r_k||/||r_{k-1}|| is
worse than `fallback_ratio` (0.5) switches the learned levels back to Chebyshev3 for the rest
of the solve (see SSGMG's `fallback_switches` and cg_add_fallback).
"""
import ctypes as C

import numpy as np

from . import FLOAT32, FLOAT64, SSHEX8
from . import (CNN, DeviceArray, SfemError, _check, _stream_ptr, cg, cg_add_fallback, gather_elements, lib,
               learned_smoother, set_constrained_value, switch)


class TrainingSet:
    """(z, target) pairs of one sshex8 level, float32 [E][channels][(L+1)^3] each, resident on the
    device."""

    def __init__(self, mesh, channels=1):
        self.mesh = mesh
        self.level = mesh.level
        self.channels = channels
        self.nxe = (mesh.level + 1) ** 3
        self.E = mesh.nelements
        self.pairs = []

    def add(self, z, t):
        self.pairs.append((z, t))

    def __len__(self):
        return len(self.pairs)


def collect_pairs(hierarchy, n_problems=4, cycles=3, seed=0, block_size=None, channels=1, first_cycle=0,
                  into=None):
    """Run plain V-cycles of `hierarchy` (an SSGMG with max_it=1) on A x = 0 from random initial
    guesses and record (D^-1 r, e) at the finest level after first_cycle..cycles cycles: for vector
    problems one pair per component (channels 1), or one pair with the components as channels
    (channels = block size, the vector network).  `into`: a TrainingSet to extend."""
    fine = hierarchy.fine
    if fine.mesh.element_type != SSHEX8:
        raise SfemError("collect_pairs: the finest level must be an sshex8 mesh")
    bs = block_size or hierarchy.block_size
    dtype = np.float64 if fine.op.real_type == FLOAT64 else np.float32
    n = fine.n
    rng = np.random.default_rng(seed)
    diag = DeviceArray(n, dtype)
    fine.A.diag(diag)
    inv = DeviceArray.from_numpy(1.0 / diag.numpy())
    zero_rhs = DeviceArray(n, dtype)
    ts = into if into is not None else TrainingSet(fine.mesh, channels)
    ne = ts.E * ts.nxe
    r = DeviceArray(n, dtype)
    for _ in range(n_problems):
        x = DeviceArray.from_numpy(rng.uniform(-1, 1, n).astype(dtype))
        set_constrained_value(fine.A, 0.0, x)
        for j in range(cycles + 1):
            if j:
                hierarchy.mg.apply(zero_rhs, x)
            if j < first_cycle:
                continue
            r.zero()
            fine.A.apply(x, r)  # r = A x = -(b - A x) with b = 0; e = -x: use z = D^-1 A x, t = x
            if channels > 1:
                z = DeviceArray(ne * bs, np.float32)
                t = DeviceArray(ne * bs, np.float32)
                gather_elements(fine.mesh, r, z, bs, -1, scale=inv)
                gather_elements(fine.mesh, x, t, bs, -1)
                ts.add(z, t)
                continue
            for c in range(bs):
                z = DeviceArray(ne, np.float32)
                t = DeviceArray(ne, np.float32)
                gather_elements(fine.mesh, r, z, bs, c, scale=inv)
                gather_elements(fine.mesh, x, t, bs, c)
                ts.add(z, t)
    return ts


def coefficient_field(seed, n_elements):
    """Per-macro-element diffusion coefficient of one variable-coefficient training or test problem:
    log-normal and piecewise constant, with log10 contrast c ~ U[0, 3] drawn from the seed:
    log10 k_e = c (xi_e - min xi) / (max xi - min xi), xi_e ~ N(0, 1), so k ranges over [1, 10^c].
    Returns (k float32 [n_elements], c)."""
    rng = np.random.default_rng(seed)
    c = float(rng.uniform(0.0, 3.0))
    xi = rng.normal(size=n_elements)
    span = float(xi.max() - xi.min())
    t = (xi - xi.min()) / span if span > 0 else np.zeros(n_elements)
    return (10.0 ** (c * t)).astype(np.float32), c


def protocol_pairs(build, seeds, elements_per_problem=None, cycles=1, channels=1, pick_seed=0):
    """Training (or held-out test) set of the variable-coefficient protocol: for each seed s, the
    plain hierarchy build(s) (an SSGMG whose coefficients come from coefficient_field(s)) runs
    `cycles` plain V-cycles on A x = 0 from a random initial guess (seeded by s), and the pair
    (D^-1 r, e) after the last one is kept.  elements_per_problem: keep only that many consecutive
    macro elements of each pair (from a seeded random start), so that thousands of problems fit
    on the device."""
    ts = None
    rng = np.random.default_rng(pick_seed)
    for s in seeds:
        h = build(s)
        full = collect_pairs(h, n_problems=1, cycles=cycles, first_cycle=cycles, seed=int(s), channels=channels)
        if elements_per_problem is None or elements_per_problem >= full.E:
            part = full
        else:
            part = TrainingSet(full.mesh, channels)
            part.E = elements_per_problem
            per = full.nxe * channels
            e0 = int(rng.integers(0, full.E - elements_per_problem + 1))
            for z, t in full.pairs:
                zs = DeviceArray(elements_per_problem * per, np.float32)
                tt = DeviceArray(elements_per_problem * per, np.float32)
                for src, dst in ((z, zs), (t, tt)):
                    _check(lib.sfem_amd_blas_copy(FLOAT32, elements_per_problem * per, _offset(src, e0, per),
                                                  C.c_void_p(dst.ptr), None), "blas_copy")
                part.add(zs, tt)
        if ts is None:
            ts = part
        else:
            ts.pairs.extend(part.pairs)
        del h
    return ts


def _offset(arr, elements_before, values_per_element):
    return C.c_void_p(arr.ptr + 4 * elements_before * values_per_element)


def train(cnn, data, steps=200, lr=2e-3, batch_elements=512, seed=0, log_every=0):
    """Adam on minibatches of `batch_elements` consecutive macro elements of random pairs.
    Returns the loss history."""
    rng = np.random.default_rng(seed)
    nb = min(batch_elements, data.E)
    hist = []
    loss = C.c_double()
    for it in range(steps):
        z, t = data.pairs[rng.integers(len(data))]
        e0 = int(rng.integers(0, data.E - nb + 1))
        per = data.nxe * data.channels
        _check(lib.sfem_amd_cnn_loss_grad(cnn.h, data.level, nb, _offset(z, e0, per), _offset(t, e0, per),
                                          C.byref(loss), None, None), "cnn_loss_grad")
        if not np.isfinite(loss.value):
            raise SfemError("train: loss is not finite")
        cnn.adam_step(lr)
        hist.append(loss.value)
        if log_every and (it + 1) % log_every == 0:
            print(f"  step {it + 1}: loss {np.mean(hist[-log_every:]):.4f}", flush=True)
    return np.array(hist)


def evaluate(cnn, data, mode=0):
    """Mean normalised loss over all pairs (whole elements), FP32 path or fused MFMA path."""
    out = DeviceArray(data.E * data.nxe * data.channels, np.float32)
    vals = []
    for z, t in data.pairs:
        cnn.forward(data.level, z, out, mode=mode)
        o, tt = out.numpy().astype(np.float64), t.numpy().astype(np.float64)
        vals.append(float(np.sum((o - tt) ** 2) / np.sum(tt ** 2)))
    return float(np.mean(vals))


def make_learned_smoother(level, cnn, steps=1, damping=1.0, mode=1, block_size=1):
    """Learned smoother for one SSGMG level (constrained operator, its diagonal); the vector network
    runs the FP32 path."""
    dtype = np.float64 if level.op.real_type == FLOAT64 else np.float32
    d = DeviceArray(level.n, dtype)
    level.A.diag(d)
    if getattr(cnn, "channels", 1) > 1:
        mode = 0
    return learned_smoother(level.A, level.mesh, cnn, d, block_size=block_size, steps=steps, damping=damping,
                            mode=mode)


def flexible_cg_with_fallback(A, precond, switches, rtol=1e-10, max_it=500, fallback_ratio=0.5):
    """FP64 flexible CG whose per-iteration residual monitor switches every op in `switches`
    to its fallback when ||r_k|| > fallback_ratio * ||r_{k-1}||."""
    solver = cg(A, precond, flexible=True, rtol=rtol, max_it=max_it)
    for sw in switches:
        cg_add_fallback(solver, sw, fallback_ratio)
    return solver
