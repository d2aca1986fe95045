#!/usr/bin/env python3
"""Benchmark: sshex8 linear elasticity, level 16 on a 24^3 macro hex8 unit cube (171,199,875 DOF,
57,066,625 nodes; Young's modulus log-uniform in [0.1, 10] per macro element, nu = 0.3).

One step = one FP64 matrix-free apply y += A x (x ~ U[-1,1], seed 0) on one MI355X.  Prints ONE JSON
line: throughput in DOF/s, the apply's roofline (algorithmic bytes over its HIP-event-timed average
duration), the FP32 apply beside it, time to true relative residual 1e-10 of FP64 MG-PCG and of the
FP32/BF16 V-cycle in FP64 flexible CG on the same workload, the same for the ~50M DOF elasticity
problem (L=8 on 32^3) and for the sshex8 Laplacian L=8 on 16^3 (its apply as well), and the CPU
baseline: the oracle's OpenMP restatement of the reference's loop, built on this host with
-O3 -fopenmp -march=native and timed on a bounded slab of the workload.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-solve] [--no-cpu] [--no-extra]
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "apply DOF/s (% HBM roofline); time-to-1e-10 MG-PCG, mixed/learned vs FP64"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure


def build_problem(n_macro=16, L=8):
    import sfem_amd as S
    mel, mpts = S.box_mesh(n_macro + 1, n_macro + 1, n_macro + 1, element_type=S.HEX8)
    el, pts = S.sshex8_refine(L, mel, mpts)
    fff = S.fff(S.SSHEX8, L, el, pts)
    return mel, mpts, el, pts, fff


def progress(msg):
    """Stage messages on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


PROFILE_LAPLACIAN = os.path.join(ROOT, "profiles", "sshex8_laplacian_apply_L8_16cube.md")
PROFILE_ELASTICITY = os.path.join(ROOT, "profiles", "sshex8_elasticity_apply_L16_24cube.md")
PROFILE_HEX8 = os.path.join(ROOT, "profiles", "hex8_elasticity_apply_128cube.md")


def profiled_traffic(path):
    """HBM bytes per apply of a workload measured by rocprofv3 --pmc (FETCH_SIZE and WRITE_SIZE,
    calibrated with a copy kernel of known size), read from the committed profile summary that
    states the commands; None without it."""
    try:
        with open(path) as f:
            m = re.search(r"Total per apply: ([0-9.]+) MB", f.read())
        return float(m.group(1)) * 1e6 if m else None
    except OSError:
        return None


def algorithmic_bytes(N, E, L, value_bytes=8, geom_bytes=4):
    # one read of x, one write of y, int32 connectivity, 6 FFF words per macro element
    return 2 * value_bytes * N + 4 * (L + 1) ** 3 * E + geom_bytes * 6 * E


def elasticity_algorithmic_bytes(N, E, L, value_bytes=8, geom_bytes=4):
    # x read + y written, int32 connectivity, 9 adjugate + 1 determinant words and 1 coefficient per
    # macro element (config 5: 2,739,198,000 + 271,669,248 + 552,960 + 55,296 = 3,011,475,504 B)
    return 2 * value_bytes * N + 4 * (L + 1) ** 3 * E + geom_bytes * 10 * E + 4 * E


def cpu_baseline(workload, size, budget_s):
    """The oracle's OpenMP restatement of the reference's CPU loop (oracle/cpu_bench.c), compiled on
    this host with the reference's release flags (-O3 -fopenmp -march=native) and run in a child
    process with OMP_PROC_BIND=close, OMP_PLACES=cores and this process's OMP_NUM_THREADS (the
    host's CPU share for one GPU, else all cores)."""
    import subprocess
    import tempfile
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    src = [os.path.join(ROOT, "oracle", "cpu_bench.c"), os.path.join(ROOT, "oracle", "sfem_oracle.c")]
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "cpu_bench")
        flags = ["-O3", "-march=native", "-fopenmp", "-std=gnu11"]
        r = subprocess.run(["gcc", *flags, "-I", os.path.join(ROOT, "oracle"), *src, "-lm", "-o", exe],
                           capture_output=True, text=True)
        if r.returncode != 0:
            return {"error": "gcc failed: " + r.stderr[-400:]}
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        r = subprocess.run([exe, workload, str(size), str(budget_s)], capture_output=True, text=True, env=env,
                           timeout=600)
        if r.returncode != 0:
            return {"error": "cpu_bench failed: " + r.stderr[-400:]}
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["build"] = "gcc " + " ".join(flags)
    out["omp"] = f"OMP_NUM_THREADS={threads} OMP_PROC_BIND=close OMP_PLACES=cores"
    return out


def solve_times(S, mesh, fff, pts, n_macro, rtol=1e-10, learned_steps=1):
    """Time to ||b - A x||/||b|| <= rtol (recomputed in FP64 with the FP64 operator) of
      fp64_eager:    FP64 MG-PCG, Chebyshev3 V-cycle, the reference's FP64 coarse PCG, every
                     kernel launched from the host;
      fp64:          the same V-cycle with a direct (dense inverse) coarse solve, replayed as a HIP
                     graph; PCG stopping on ||r||/||b||;
      mixed:         that V-cycle in FP32 with BF16 geometry, graph-replayed, in FP64 flexible CG;
      mixed_learned: as mixed, with the learned CNN smoother on the finest level (trained at
                     setup on V-cycle pairs; Chebyshev3 fallback when an iteration reduces the
                     residual by less than 2x); `learned_steps` CNN steps per smoothing (one is
                     faster to 1e-10 than two on the 16^3, L=8 problem: 7 iterations at half the
                     inference cost against 6).
    Every timed solve starts from x = 0 after one untimed warm-up solve."""
    from sfem_amd import gmg, learned
    N = pts.shape[1]
    bc = np.where((pts.min(0) < 1e-6) | (pts.max(0) > 1 - 1e-6))[0].astype(np.int32)
    rhs = np.random.default_rng(1).uniform(0, 1, N)
    rhs[bc] = 0
    b = S.DeviceArray.from_numpy(rhs)
    out = {}

    def true_res(A, x):
        r = S.DeviceArray(N)
        A.apply(x, r)
        return float(np.linalg.norm(rhs - r.numpy()) / np.linalg.norm(rhs))

    box = (n_macro, 1.0)
    hierarchy = "V-cycle 8-4-2-1 | macro 16^3 -> 8-4-2-1 super levels, Chebyshev3 degree 3"
    t0 = time.perf_counter()
    h_ref = gmg.SSGMG(mesh, "laplacian", fff, bc, macro_box=box)
    setup_ref = time.perf_counter() - t0
    A64 = h_ref.fine.A
    h32 = None
    for name in ("fp64_eager", "fp64", "fp64_power10", "mixed", "mixed_learned"):
        progress(f"laplacian solve: {name}")
        switches = []
        extra = {}
        t0 = time.perf_counter()
        if name == "fp64_eager":
            solver = S.cg(A64, h_ref.mg, rtol=rtol, max_it=500, stop_on_residual=True)
            setup = setup_ref
            desc = hierarchy + ", FP64; FP64 coarse PCG (host-synchronised); PCG stopping on ||r||/||b||"
        elif name == "fp64":
            h64 = gmg.SSGMG(mesh, "laplacian", fff, bc, macro_box=box, coarse="dense")
            solver = S.cg(A64, h64.preconditioner(graph=True), rtol=rtol, max_it=500, stop_on_residual=True)
            setup = time.perf_counter() - t0
            desc = hierarchy + ", FP64; dense direct coarse solve; V-cycle replayed as a HIP graph"
        elif name == "fp64_power10":
            # the workload's literal "eigen-bound from 10 power iterations" (11: the reference rounds
            # up to odd): lambda_max comes out low and the V-cycle loses its efficiency
            hp = gmg.SSGMG(mesh, "laplacian", fff, bc, macro_box=box, coarse="dense", eig_its=10, eig_tol=0.0)
            solver = S.cg(A64, hp.preconditioner(graph=True), rtol=rtol, max_it=500, stop_on_residual=True)
            setup = time.perf_counter() - t0
            desc = hierarchy + ", FP64, 10 power iterations per bound; dense coarse solve; HIP graph"
        elif name == "mixed":
            h32 = gmg.SSGMG(mesh, "laplacian", fff, bc, real_type=S.FLOAT32, macro_box=box, coarse="dense")
            solver = S.cg(A64, h32.preconditioner(graph=True), flexible=True, rtol=rtol, max_it=500)
            setup = time.perf_counter() - t0
            desc = hierarchy + " in FP32 with BF16 geometry, dense coarse solve, HIP graph; outer FP64 flexible CG"
        else:
            data = learned.collect_pairs(h32, n_problems=4, cycles=2, seed=0)
            cnn = S.CNN(16, 4, seed=0)
            hist = learned.train(cnn, data, steps=300, lr=2e-3, batch_elements=16, seed=0)
            t_train = time.perf_counter() - t0
            hl = gmg.SSGMG(mesh, "laplacian", fff, bc, real_type=S.FLOAT32, macro_box=box, coarse="dense",
                           learned=cnn, learned_steps=learned_steps)
            switches = hl.switches
            solver = learned.flexible_cg_with_fallback(A64, hl.preconditioner(graph=True), switches, rtol=rtol,
                                                       max_it=500)
            setup = time.perf_counter() - t0
            extra = {"training_seconds": t_train, "training_loss_last20": float(np.mean(hist[-20:])),
                     "learned_steps_per_smoothing": learned_steps}
            desc = ("mixed, with the finest-level Chebyshev3 replaced by the learned CNN smoother (BF16 MFMA "
                    "inference); Chebyshev3 fallback at per-iteration reduction > 0.5")
        # one untimed solve to warm up code paths (and capture graphs), then the timed one from x = 0
        x = S.DeviceArray(N)
        solver.apply(b, x)
        for sw in switches:
            S.switch_select(sw, 0)
        ev0 = sum(S.switch_state(sw)[1] for sw in switches)
        x.zero()
        S.synchronize()
        t0 = time.perf_counter()
        solver.apply(b, x)
        S.synchronize()
        t = time.perf_counter() - t0
        out[name] = {"seconds": t, "iterations": solver.iterations, "status": solver.last_status,
                     "true_rel_residual": true_res(A64, x), "setup_seconds": setup,
                     "preconditioner": desc, **extra}
        if switches:
            out[name]["fallback_events"] = int(sum(S.switch_state(sw)[1] for sw in switches) - ev0)
    for name in ("fp64", "fp64_power10", "mixed", "mixed_learned"):
        out[f"speedup_{name}_vs_fp64_eager"] = out["fp64_eager"]["seconds"] / out[name]["seconds"]
    out["speedup_mixed_vs_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["speedup_mixed_learned_vs_fp64"] = out["fp64"]["seconds"] / out["mixed_learned"]["seconds"]
    return out


def elasticity_solve_times(S, n_macro=32, L=8, rtol=1e-10, with_learned=False, coeff_seed=None):
    """sshex8 linear elasticity, level L on an n_macro^3 macro hex8 unit cube (3 (nL+1)^3 DOF:
    50,923,779 at L=8, n=32; 171,199,875 at L=16, n=24), nu = 0.3, clamped at z = 0, body force
    (0, 0, -1); E = 1, or with coeff_seed a per-macro-element Young's modulus log-uniform in
    [0.1, 10].  Mesh, refined points and adjugates are built on the device.  Time to true relative
    residual rtol of FP64 MG-PCG (graph-replayed V-cycle, dense coarse solve) and of the FP32/BF16
    V-cycle in FP64 flexible CG (optionally with the learned smoother per component)."""
    from sfem_amd import gmg, learned
    t0 = time.perf_counter()
    mel, mpts = S.box_mesh(n_macro + 1, n_macro + 1, n_macro + 1, element_type=S.HEX8)
    mesh, pts_d = S.sshex8_refine_device(L, mel, mpts)
    adj, det = S.mesh_adjugate_det(mesh, pts_d)
    N = mesh.nnodes
    bc = np.where(pts_d.numpy().reshape(3, N)[2] < 1e-6)[0].astype(np.int32)
    del pts_d
    coeff = None
    if coeff_seed is not None:
        coeff = (10.0 ** np.random.default_rng(coeff_seed).uniform(-1, 1, n_macro ** 3)).astype(np.float32)
    nu = 0.3
    mu, lam = 1.0 / (2 * (1 + nu)), nu / ((1 + nu) * (1 - 2 * nu))
    rhs = np.zeros((N, 3))
    rhs[:, 2] = -1
    rhs[bc] = 0
    rhs = rhs.ravel()
    b = S.DeviceArray.from_numpy(rhs)
    mesh_s = time.perf_counter() - t0
    box = (n_macro, 1.0)
    out = {"workload": f"sshex8 linear elasticity L={L} on a {n_macro}^3 macro cube, {3 * N} DOF, nu=0.3, "
                       + ("E log-uniform in [0.1,10] per macro element (seed %d), " % coeff_seed
                          if coeff_seed is not None else "E=1, ")
                       + "clamped z=0, body force (0,0,-1)", "dofs": 3 * N, "mesh_seconds": mesh_s}
    # jumping coefficients: Chebyshev3 on D^-1 A (111 against 23 FP64 iterations at L=16, n=24)
    smoother = "cheb3" if coeff is None else "cheb3_jacobi"
    out["smoother"] = smoother
    # Chebyshev bounds from Lanczos (at most 200 steps, 1e-5 relative): 45 fine-level applies at
    # L=16 where the power method takes ~390, and closer to lambda_max (it stops ~1.7% low here)
    common = dict(mu=mu, lam=lam, macro_box=box, coarse="dense", coeff=coeff, smoother=smoother, lanczos_steps=200)
    out["eigen_bound"] = "Lanczos, at most 200 steps, relative tolerance 1e-5 (power method: sfem_amd.gmg default)"
    progress(f"elasticity: {3 * N} DOF mesh ready in {mesh_s:.1f}s; building FP64 hierarchy")
    t0 = time.perf_counter()
    h64 = gmg.SSGMG(mesh, "elasticity", (adj, det), bc, **common)
    A64 = h64.fine.A
    setups = {"fp64": time.perf_counter() - t0}
    cases = [("fp64", h64, S.FLOAT64)]
    progress("elasticity: building FP32 hierarchy")
    t0 = time.perf_counter()
    h32 = gmg.SSGMG(mesh, "elasticity", (adj, det), bc, real_type=S.FLOAT32, **common)
    setups["mixed"] = time.perf_counter() - t0
    cases.append(("mixed", h32, S.FLOAT32))
    if with_learned:
        # the 3-channel network acts on the finest level with micro-grid level <= 8; its training pairs
        # come from the plain FP32 V-cycle started on that level
        progress("elasticity: training the learned smoother (3-channel network)")
        t0 = time.perf_counter()
        if L > 8:
            lm = mesh.derefine(8)
            ht = gmg.SSGMG(lm, "elasticity", (adj, det), bc[bc < lm.nnodes], real_type=S.FLOAT32, **common)
        else:
            ht = h32
        data = learned.collect_pairs(ht, n_problems=2, cycles=2, seed=0, channels=3)
        del ht
        cnn = S.CNN(16, 4, seed=0, channels=3)
        hist = learned.train(cnn, data, steps=300, lr=2e-3, batch_elements=16, seed=0)
        hl = gmg.SSGMG(mesh, "elasticity", (adj, det), bc, real_type=S.FLOAT32, learned=cnn, learned_steps=1,
                       **common)
        setups["mixed_learned"] = time.perf_counter() - t0
        out["learned_training_loss_last20"] = float(np.mean(hist[-20:]))
        cases.append(("mixed_learned", hl, S.FLOAT32))

    def true_res(x):
        r = S.DeviceArray(3 * N)
        A64.apply(x, r)
        return float(np.linalg.norm(rhs - r.numpy()) / np.linalg.norm(rhs))

    for name, h, rt in cases:
        progress(f"elasticity solve: {name}")
        P = h.preconditioner(graph=True)
        if rt == S.FLOAT64:
            solver = S.cg(A64, P, rtol=rtol, max_it=1000, stop_on_residual=True)
        else:
            solver = learned.flexible_cg_with_fallback(A64, P, h.switches, rtol=rtol, max_it=1000)
        x = S.DeviceArray(3 * N)
        solver.apply(b, x)  # warm-up (captures the graph)
        h.select_smoothers(0)
        x.zero()
        S.synchronize()
        t0 = time.perf_counter()
        solver.apply(b, x)
        S.synchronize()
        t = time.perf_counter() - t0
        out[name] = {"seconds": t, "iterations": solver.iterations, "status": solver.last_status,
                     "true_rel_residual": true_res(x), "setup_seconds": setups[name]}
        if h.switches:
            out[name]["fallback_active"] = int(S.switch_state(h.switches[0])[0])
            out[name]["fallback_events"] = int(S.switch_state(h.switches[0])[1])
    out["speedup_mixed_vs_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    if with_learned:
        out["speedup_mixed_learned_vs_fp64"] = out["fp64"]["seconds"] / out["mixed_learned"]["seconds"]
    return out


def config5_apply(S, steps, warmup):
    """The headline workload: FP64 (and, beside it, FP32) sshex8 elasticity apply, L=16 on 24^3."""
    n_macro, L = 24, 16
    progress("config 5: refining the 24^3 macro cube to level 16 on the device")
    mel, mpts = S.box_mesh(n_macro + 1, n_macro + 1, n_macro + 1, element_type=S.HEX8)
    mesh, pts_d = S.sshex8_refine_device(L, mel, mpts)
    adj, det = S.mesh_adjugate_det(mesh, pts_d)
    del pts_d
    N, E = 3 * mesh.nnodes, n_macro ** 3
    coeff = (10.0 ** np.random.default_rng(0).uniform(-1, 1, E)).astype(np.float32)
    nu = 0.3
    mu, lam = 1.0 / (2 * (1 + nu)), nu / ((1 + nu) * (1 - 2 * nu))
    op = S.linear_elasticity(mesh, adj, det, mu, lam, coeff)
    x = S.DeviceArray.from_numpy(np.random.default_rng(0).uniform(-1, 1, N))
    y = S.DeviceArray(N)
    progress(f"config 5 apply benchmark: {steps} steps")
    total_ms = S.benchmark(op, x, y, warmup=warmup, reps=steps)
    out = {"N": N, "E": E, "L": L, "nodes": mesh.nnodes, "total_ms": total_ms}
    del op, x, y
    op32 = S.linear_elasticity(mesh, adj, det, mu, lam, coeff, real_type=S.FLOAT32)
    x32 = S.DeviceArray.from_numpy(np.random.default_rng(0).uniform(-1, 1, N).astype(np.float32), np.float32)
    y32 = S.DeviceArray(N, np.float32)
    ms32 = S.benchmark(op32, x32, y32, warmup=warmup, reps=steps) / steps
    B32 = elasticity_algorithmic_bytes(N, E, L, value_bytes=4, geom_bytes=2)
    out["fp32"] = {"ms_per_apply": ms32, "dof_per_s": N / (ms32 * 1e-3), "algorithmic_bytes": B32,
                   "GB_per_s": B32 / (ms32 * 1e-3) / 1e9,
                   "hbm_frac": B32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "form": "FP32 vectors, BF16 geometry, element form on the matrix cores (bf16x3 products, "
                           "FP32 accumulation)"}
    return out


def hex8_elasticity_config3(S, args, n=128, rtol=1e-10):
    """hex8 linear elasticity on a 128^3 hex8 unit cube (6,440,067 DOF): E = 1, nu = 0.3, clamped at
    z = 0, body force (0, 0, -1).  The FP64 apply and the FP32 one (bf16 adjugate and determinant words)
    on element patches, the FP32 apply's error against the FP64 one, and the time to true relative
    residual rtol of FP64 MG-PCG and of the FP32 hierarchy in FP64 flexible CG.  The multigrid
    continues below the hex8 grid on the super-mesh levels 8, 4, 2, 1 of an 8^3 box (graph-replayed
    V-cycles, dense coarse solve)."""
    from sfem_amd import gmg, learned
    progress(f"config 3: hex8 {n}^3 mesh")
    t0 = time.perf_counter()
    el, pts = S.box_mesh(n + 1, n + 1, n + 1, element_type=S.HEX8)
    adj, det = S.adjugate_det(S.HEX8, 1, el, pts)
    nn, E = pts.shape[1], el.shape[1]
    N = 3 * nn
    mesh = S.Mesh(S.HEX8, 1, el, nn)
    bc = np.where(pts[2] < 1e-6)[0].astype(np.int32)
    nu = 0.3
    mu, lam = 1.0 / (2 * (1 + nu)), nu / ((1 + nu) * (1 - 2 * nu))
    mesh_s = time.perf_counter() - t0
    out = {"workload": f"hex8 linear elasticity on a {n}^3 hex8 unit cube ({N} DOF), E=1, nu=0.3, clamped z=0, "
                       "body force (0,0,-1)", "dofs": N, "elements": E, "mesh_seconds": mesh_s}
    x64 = S.DeviceArray.from_numpy(np.random.default_rng(0).uniform(-1, 1, N))
    x32 = S.DeviceArray.from_numpy(x64.numpy().astype(np.float32), np.float32)
    ys = {}
    for name, rt, x, vb, gb in (("fp64", S.FLOAT64, x64, 8, 4), ("fp32", S.FLOAT32, x32, 4, 2)):
        op = S.linear_elasticity(mesh, adj, det, mu, lam, real_type=rt)
        y = S.DeviceArray(N, x.dtype)
        t0 = time.perf_counter()
        op.apply(x, y)  # first use builds the element patches
        S.synchronize()
        first = time.perf_counter() - t0
        y.zero()
        op.apply(x, y)
        S.synchronize()
        ys[name] = y.numpy().astype(np.float64)
        reps = max(args.steps, 20)
        ms = S.benchmark(op, x, y, warmup=3, reps=reps) / reps
        B = 2 * vb * N + 4 * 8 * E + gb * 10 * E
        out["apply_" + name] = {"ms_per_apply": ms, "dof_per_s": N / (ms * 1e-3), "algorithmic_bytes": B,
                                "GB_per_s": B / (ms * 1e-3) / 1e9, "hbm_frac": B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                "first_apply_seconds": first}
    out["apply_fp64"]["traffic"] = profiled_traffic(PROFILE_HEX8)
    out["apply_fp64"]["traffic_source"] = ("profiles/hex8_elasticity_apply_128cube.md (rocprofv3 --pmc FETCH_SIZE, "
                                           "WRITE_SIZE per dispatch, patch kernel + gather, copy-calibrated)")
    out["apply_fp32"]["rel_error_vs_fp64"] = float(np.linalg.norm(ys["fp32"] - ys["fp64"]) /
                                                   np.linalg.norm(ys["fp64"]))
    out["apply_fp32"]["note"] = ("FP32 sum-factorised arithmetic on bf16 geometry words; error against the FP64 apply, "
                                 "itself within 1e-12 of the oracle (tests/test_gpu_ops.py)")
    if args.no_solve:
        return out
    rhs = np.zeros((nn, 3))
    rhs[:, 2] = -1
    rhs[bc] = 0
    rhs = rhs.ravel()
    b = S.DeviceArray.from_numpy(rhs)
    common = dict(mu=mu, lam=lam, macro_box=(n, 1.0), coarse="dense")
    progress("config 3: FP64 and FP32 hierarchies")
    t0 = time.perf_counter()
    h64 = gmg.SSGMG(mesh, "elasticity", (adj, det), bc, **common)
    setups = {"fp64": time.perf_counter() - t0}
    t0 = time.perf_counter()
    h32 = gmg.SSGMG(mesh, "elasticity", (adj, det), bc, real_type=S.FLOAT32, **common)
    setups["mixed"] = time.perf_counter() - t0
    A64 = h64.fine.A
    out["levels"] = [f"{'hex8' if lv.mesh.element_type == S.HEX8 else 'sshex8'} L={lv.mesh.level} "
                     f"({3 * lv.mesh.nnodes} DOF)" for lv in h64.levels]
    for name, h, rt in (("fp64", h64, S.FLOAT64), ("mixed", h32, S.FLOAT32)):
        progress(f"config 3 solve: {name}")
        P = h.preconditioner(graph=True)
        if rt == S.FLOAT64:
            solver = S.cg(A64, P, rtol=rtol, max_it=1000, stop_on_residual=True)
        else:
            solver = learned.flexible_cg_with_fallback(A64, P, h.switches, rtol=rtol, max_it=1000)
        x = S.DeviceArray(N)
        solver.apply(b, x)  # warm-up (captures the graph)
        x.zero()
        S.synchronize()
        t0 = time.perf_counter()
        solver.apply(b, x)
        S.synchronize()
        t = time.perf_counter() - t0
        r = S.DeviceArray(N)
        A64.apply(x, r)
        res = float(np.linalg.norm(rhs - r.numpy()) / np.linalg.norm(rhs))
        out["solve_" + name] = {"seconds": t, "iterations": solver.iterations, "status": solver.last_status,
                                "true_rel_residual": res, "setup_seconds": setups[name]}
    out["speedup_mixed_vs_fp64"] = out["solve_fp64"]["seconds"] / out["solve_mixed"]["seconds"]
    return out


def learned_config4(S, args, n_macro=32, L=8, n_train=None, train_macro=8, held_out=(1000, 1001), rtol=1e-10):
    """Variable-coefficient sshex8 Poisson, level 8 on a 32^3 macro unit cube (16,974,593 DOF), zero
    Dirichlet on all faces, RHS U[0,1]; per-macro-element coefficients log-normal with log10 contrast
    U[0,3] (learned.coefficient_field).  The learned-smoother protocol: training pairs (D^-1 r, e)
    after one plain V-cycle (FP32 hierarchy, Chebyshev3 on D^-1 A) on n_train seeded problems of an
    8^3-macro copy (the network acts per macro element, so its training problems need not be the full
    mesh; the full protocol takes 2,000), Adam on the project's kernels; then on held-out seeds at full
    size, the time to true relative residual rtol of FP64 MG-PCG, of the FP32 hierarchy in FP64
    flexible CG, and of the latter with the trained network (--learned-steps per smoothing) on the
    finest level.  Every
    hierarchy continues below the macro mesh (super levels), dense coarse solve, graph-replayed
    V-cycles."""
    from sfem_amd import gmg, learned
    if n_train is None:
        n_train = args.protocol_problems
    adam_steps = args.adam_steps
    progress(f"config 4: training set ({n_train} problems on a {train_macro}^3 macro copy)")
    t0 = time.perf_counter()
    tmel, tmpts = S.box_mesh(train_macro + 1, train_macro + 1, train_macro + 1, element_type=S.HEX8)
    tel, tpts = S.sshex8_refine(L, tmel, tmpts)
    tfff = S.fff(S.HEX8, 1, tmel, tmpts)
    tmesh = S.Mesh(S.SSHEX8, L, tel, tpts.shape[1])
    tbc = np.where((tpts.min(0) < 1e-6) | (tpts.max(0) > 1 - 1e-6))[0].astype(np.int32)

    def build_train(seed):
        k, _ = learned.coefficient_field(seed, train_macro ** 3)
        return gmg.SSGMG(tmesh, "laplacian", tfff, tbc, coeff=k, real_type=S.FLOAT32, smoother="cheb3_jacobi",
                         macro_box=(train_macro, 1.0))

    data = None
    for c0 in range(0, n_train, 250):  # in chunks, with a progress line each
        part = learned.protocol_pairs(build_train, range(c0, min(n_train, c0 + 250)), cycles=1)
        if data is None:
            data = part
        else:
            data.pairs.extend(part.pairs)
        if n_train > 250:
            progress(f"config 4: pairs of {min(n_train, c0 + 250)} problems")
    t_pairs = time.perf_counter() - t0
    cnn = S.CNN(16, 4, seed=0)
    hist = learned.train(cnn, data, steps=adam_steps, lr=2e-3, batch_elements=32, seed=0)
    t_train = time.perf_counter() - t0 - t_pairs
    test_small = learned.protocol_pairs(build_train, [s + 5000 for s in held_out], cycles=1)
    out = {"workload": f"sshex8 Poisson L={L} on a {n_macro}^3 macro cube, log-normal per-macro-element coefficient "
                       "(log10 contrast U[0,3]), zero Dirichlet on all faces, RHS U[0,1] (seeded)",
           "training": {"problems": n_train, "of_protocol": 2000, "train_mesh": f"{train_macro}^3 macro, L={L}",
                        "pairs": len(data), "pairs_seconds": t_pairs, "train_seconds": t_train, "adam_steps": adam_steps,
                        "learned_steps_per_smoothing": args.learned_steps,
                        "loss_last20": float(np.mean(hist[-20:])),
                        "held_out_loss": learned.evaluate(cnn, test_small)}}
    del data, test_small
    progress(f"config 4: {n_macro}^3 macro mesh")
    mel, mpts = S.box_mesh(n_macro + 1, n_macro + 1, n_macro + 1, element_type=S.HEX8)
    mesh, pts_d = S.sshex8_refine_device(L, mel, mpts)
    fff = S.fff(S.HEX8, 1, mel, mpts)
    N = mesh.nnodes
    p = pts_d.numpy().reshape(3, N)
    bc = np.where((p.min(0) < 1e-6) | (p.max(0) > 1 - 1e-6))[0].astype(np.int32)
    del pts_d, p
    out["dofs"] = N
    box = (n_macro, 1.0)
    rows = []
    for seed in held_out:
        k, c = learned.coefficient_field(seed, n_macro ** 3)
        rhs = np.random.default_rng(seed).uniform(0, 1, N)
        rhs[bc] = 0
        b = S.DeviceArray.from_numpy(rhs)
        common = dict(coeff=k, smoother="cheb3_jacobi", macro_box=box, coarse="dense")
        row = {"seed": seed, "log10_contrast": c}
        h64 = A64 = None
        for name in ("fp64", "mixed", "mixed_learned"):
            progress(f"config 4 seed {seed}: {name}")
            t0 = time.perf_counter()
            if name == "fp64":
                h = h64 = gmg.SSGMG(mesh, "laplacian", fff, bc, **common)
                A64 = h64.fine.A
                solver = S.cg(A64, h.preconditioner(graph=True), rtol=rtol, max_it=500, stop_on_residual=True)
            elif name == "mixed":
                h = gmg.SSGMG(mesh, "laplacian", fff, bc, real_type=S.FLOAT32, **common)
                solver = S.cg(A64, h.preconditioner(graph=True), flexible=True, rtol=rtol, max_it=500)
            else:
                h = gmg.SSGMG(mesh, "laplacian", fff, bc, real_type=S.FLOAT32, learned=cnn,
                              learned_steps=args.learned_steps, **common)
                solver = learned.flexible_cg_with_fallback(A64, h.preconditioner(graph=True), h.switches, rtol=rtol,
                                                           max_it=500)
            setup = time.perf_counter() - t0
            x = S.DeviceArray(N)
            solver.apply(b, x)  # warm-up (captures the graph)
            h.select_smoothers(0)
            ev0 = sum(S.switch_state(sw)[1] for sw in h.switches)
            x.zero()
            S.synchronize()
            t0 = time.perf_counter()
            solver.apply(b, x)
            S.synchronize()
            t = time.perf_counter() - t0
            r = S.DeviceArray(N)
            A64.apply(x, r)
            res = float(np.linalg.norm(rhs - r.numpy()) / np.linalg.norm(rhs))
            row[name] = {"seconds": t, "iterations": solver.iterations, "status": solver.last_status,
                         "true_rel_residual": res, "setup_seconds": setup}
            if h.switches:
                row[name]["fallback_events"] = int(sum(S.switch_state(sw)[1] for sw in h.switches) - ev0)
            del solver, h
        del h64, A64
        row["speedup_mixed_vs_fp64"] = row["fp64"]["seconds"] / row["mixed"]["seconds"]
        row["speedup_mixed_learned_vs_fp64"] = row["fp64"]["seconds"] / row["mixed_learned"]["seconds"]
        rows.append(row)
    out["held_out"] = rows
    out["mean_speedup_mixed_learned_vs_fp64"] = float(np.mean([r["speedup_mixed_learned_vs_fp64"] for r in rows]))
    out["mean_speedup_mixed_vs_fp64"] = float(np.mean([r["speedup_mixed_vs_fp64"] for r in rows]))
    return out


def laplacian_apply_only(S, args):
    """sshex8 Laplacian apply alone, level 8 on an n^3 macro cube (n = 16: the 2,146,689-DOF
    workload; n = 32: 16,974,593 DOF, the variable-coefficient workload's size), FP64 and FP32/BF16,
    mesh and FFF built on the device: a short run for the profiler."""
    L, n = 8, args.n_macro
    mel, mpts = S.box_mesh(n + 1, n + 1, n + 1, element_type=S.HEX8)
    mesh, pts_d = S.sshex8_refine_device(L, mel, mpts)
    fff = S.mesh_fff(mesh, pts_d)
    N, E = mesh.nnodes, n ** 3
    out = {"workload": f"sshex8 Laplacian apply y += A x, level 8 on a {n}^3 macro hex8 unit cube ({N} DOF)",
           "dofs": N}
    for name, rt, dt, vb in (("fp64", S.FLOAT64, np.float64, 8), ("fp32", S.FLOAT32, np.float32, 4)):
        op = S.laplacian(mesh, fff, real_type=rt)
        x = S.DeviceArray.from_numpy(np.random.default_rng(0).uniform(-1, 1, N).astype(dt))
        y = S.DeviceArray(N, dt)
        progress(f"laplacian apply {name}: {args.steps} steps")
        ms = S.benchmark(op, x, y, warmup=args.warmup, reps=args.steps) / args.steps
        B = algorithmic_bytes(N, E, L, value_bytes=vb, geom_bytes=4 if vb == 8 else 2)
        out[name] = {"ms_per_apply": ms, "dof_per_s": N / (ms * 1e-3), "algorithmic_bytes": B,
                     "GB_per_s": B / (ms * 1e-3) / 1e9, "hbm_frac": B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    return out


def laplacian_config2(S, args):
    """sshex8 Laplacian L=8 on a 16^3 macro cube (2,146,689 DOF): apply (hot and cold caches, and the
    atomic skeleton option) and the MG-PCG solves."""
    L, n_macro = 8, 16
    mel, mpts, el, pts, fff = build_problem(n_macro, L)
    N, E = pts.shape[1], el.shape[1]
    mesh = S.Mesh(S.SSHEX8, L, el, N)
    op = S.laplacian(mesh, fff)
    x = S.DeviceArray.from_numpy(np.random.default_rng(0).uniform(-1, 1, N))
    y = S.DeviceArray(N)
    progress("laplacian config 2 apply")
    steps = 200
    ms = S.benchmark(op, x, y, warmup=10, reps=steps) / steps
    B = algorithmic_bytes(N, E, L)
    ach = B / (ms * 1e-3) / 1e9
    out = {"workload": f"sshex8 Laplacian apply, level 8 on a 16^3 macro hex8 unit cube ({N} DOF), FP64",
           "value": N / (ms * 1e-3), "unit": "DOF/s", "ms_per_apply": ms,
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": profiled_traffic(PROFILE_LAPLACIAN),
                        "algorithmic_bytes_per_apply": B,
                        "note": "working set 46 MB fits the 256 MB MALL: back-to-back applies are cache-hot"}}
    if not args.no_flush:
        cold = S.benchmark(op, x, y, warmup=1, reps=20, flush_bytes=1 << 30) / 20
        a2 = B / (cold * 1e-3) / 1e9
        out["roofline_cold"] = {"achieved": a2, "frac": a2 / HBM_PEAK_GBS, "ms_per_apply": cold,
                                "note": "1 GiB written between applies (outside the timed events)"}
    S.set_option("atomic_skeleton", 1)
    try:
        ta = S.benchmark(op, x, y, warmup=10, reps=steps) / steps
    finally:
        S.set_option("atomic_skeleton", 0)
    out["apply_atomic_skeleton"] = {"ms_per_apply": ta, "value": N / (ta * 1e-3),
                                    "note": "FP64 atomics for the macro-element boundary sums (not reproducible)"}
    if not args.no_solve:
        out["solve"] = solve_times(S, mesh, fff, pts, n_macro)
    if not args.no_cpu:
        progress("cpu baseline (laplacian)")
        cb = cpu_baseline("laplacian", n_macro, 8.0)
        if "error" not in cb:
            out["cpu_baseline"] = {"value": cb["dof_per_s"], "unit": "DOF/s", "cores": cb["threads"], "kind": "port",
                                   "sample": "full apply of this workload, median of %d after 2 warm-up: the oracle's "
                                             "restatement of affine_sshex8_laplacian_stencil_apply_fff (stencil + "
                                             "surface stencil, atomic scatter)" % cb["reps"], **cb}
        else:
            out["cpu_baseline"] = cb
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-flush", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the headline apply (and its solves unless --no-solve): skip the ~50M DOF "
                         "elasticity solve and the Laplacian workload")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned smoother (3-channel network) in the elasticity solves")
    ap.add_argument("--protocol-problems", type=int, default=32,
                    help="learned-smoother protocol: seeded training problems (the full protocol: 2000)")
    ap.add_argument("--adam-steps", type=int, default=600, help="learned-smoother protocol: Adam steps")
    ap.add_argument("--learned-steps", type=int, default=2,
                    help="learned-smoother protocol: CNN steps per smoothing in the held-out solves")
    ap.add_argument("--n-macro", type=int, default=16,
                    help="laplacian-apply: macro elements per direction of the level-8 cube")
    ap.add_argument("--case", choices=["l8", "l16"], default="l16",
                    help="elasticity-solve: L=8 on 32^3 (E=1) or L=16 on 24^3 (E log-uniform)")
    ap.add_argument("--mode", choices=["apply", "copy", "elasticity-apply", "laplacian-apply", "elasticity-solve",
                                       "laplacian-solve", "hex8-elasticity", "learned"],
                    default="apply",
                    help="copy: stream-copy 2 GiB of FP64 per step (HBM counter calibration for rocprofv3 --pmc); "
                         "elasticity-apply: only the headline apply (a short run for the profiler); "
                         "laplacian-apply: only the sshex8 Laplacian apply on --n-macro^3 (FP64 and FP32); "
                         "elasticity-solve: only the elasticity MG-PCG solves of --case; "
                         "laplacian-solve: only the MG-PCG solves of the 2.15M-DOF Laplacian workload; "
                         "hex8-elasticity: only the 128^3 hex8 elasticity applies and solves; "
                         "learned: only the variable-coefficient learned-smoother protocol")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (this build runs on a single "
              "GPU; the workload measured is one GPU's shard, and no multi-GPU path is built).", file=sys.stderr)
        sys.exit(2)

    import sfem_amd as S
    if S.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.mode == "copy":
        n = (2 << 30) // 8  # 2 GiB per array: far beyond the 256 MB MALL
        a = S.DeviceArray.from_numpy(np.ones(n))
        b = S.DeviceArray(n)
        S.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            S.lib.sfem_amd_blas_copy(S.FLOAT64, n, S._dptr(a), S._dptr(b), None)
        S.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        print(json.dumps({"mode": "copy", "bytes_per_step": 16 * n, "GB/s": 16 * n / dt / 1e9}))
        return

    if args.mode == "learned":
        print(json.dumps({"mode": "learned", **learned_config4(S, args)}))
        return
    if args.mode == "elasticity-solve":
        kw = dict(n_macro=32, L=8) if args.case == "l8" else dict(n_macro=24, L=16, coeff_seed=0)
        print(json.dumps({"mode": "elasticity-solve",
                          **elasticity_solve_times(S, with_learned=not args.no_learned, **kw)}))
        return
    if args.mode == "laplacian-solve":
        mel, mpts, el, pts, fff = build_problem(16, 8)
        mesh = S.Mesh(S.SSHEX8, 8, el, pts.shape[1])
        print(json.dumps({"mode": "laplacian-solve", **solve_times(S, mesh, fff, pts, 16)}))
        return
    if args.mode == "laplacian-apply":
        print(json.dumps({"mode": "laplacian-apply", **laplacian_apply_only(S, args)}))
        return
    if args.mode == "hex8-elasticity":
        print(json.dumps({"mode": "hex8-elasticity", **hex8_elasticity_config3(S, args)}))
        return
    if args.mode == "elasticity-apply":
        args.no_solve = args.no_extra = args.no_cpu = True
    h = config5_apply(S, args.steps, args.warmup)
    N, E, L = h["N"], h["E"], h["L"]
    ms = h["total_ms"] / args.steps
    B = elasticity_algorithmic_bytes(N, E, L)
    achieved = B / (ms * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": N * args.steps / (h["total_ms"] * 1e-3),
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "sshex8 linear elasticity apply y += A x, level 16 on a 24^3 macro hex8 unit cube "
                               f"({N} DOF, {h['nodes']} nodes, {E} macro elements), Young's modulus log-uniform in "
                               "[0.1, 10] per macro element (seed 0), nu 0.3; FP64 vectors, float32 adjugate and "
                               "determinant, int32 ids",
                   "dofs": N, "nodes": h["nodes"], "macro_elements": E, "level": L,
                   "form": "stencil form (27-point 3x3-block stencil per macro element, DPP-broadcast coefficients) "
                           "+ fixed-order skeleton gather"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(PROFILE_ELASTICITY),
                     "traffic_source": "profiles/sshex8_elasticity_apply_L16_24cube.md (rocprofv3 --pmc FETCH_SIZE, "
                                       "WRITE_SIZE per dispatch, both kernels of one apply, copy-calibrated)",
                     "algorithmic_bytes_per_apply": B,
                     "note": "achieved over the whole apply (stencil kernel + skeleton gather), HIP events on the "
                             "launch stream; the 3.0 GB working set exceeds the 256 MB MALL"},
        "apply_fp32": h["fp32"],
    }
    # the small Laplacian workload first: its cache-resident apply is sensitive to clocks and
    # memory state after the large solves
    if not args.no_extra:
        result["laplacian_config2"] = laplacian_config2(S, args)
        # the Laplacian apply at the variable-coefficient workload's size (beyond the MALL)
        result["laplacian_apply_32cube"] = laplacian_apply_only(S, argparse.Namespace(**{**vars(args), "n_macro": 32}))
        result["hex8_elasticity_config3"] = hex8_elasticity_config3(S, args)
        if not args.no_solve:
            result["learned_config4"] = learned_config4(S, args)
    if not args.no_solve:
        result["solve_config5"] = elasticity_solve_times(S, n_macro=24, L=16, coeff_seed=0,
                                                         with_learned=not args.no_learned)
        if not args.no_extra:
            result["solve_elasticity_l8"] = elasticity_solve_times(S, with_learned=not args.no_learned)
    if not args.no_cpu:
        progress("cpu baseline (elasticity slab)")
        cb = cpu_baseline("elasticity", 2, 10.0)
        if "error" in cb:
            result["cpu_baseline"] = {"value": None, "unit": "DOF/s", "cores": None, "kind": "port", **cb}
        else:
            result["cpu_baseline"] = {
                "value": cb["dof_per_s"], "unit": "DOF/s", "cores": cb["threads"], "kind": "port",
                "sample": "full applies of a 24 x 24 x 2 slab of the 24^3 macro cube at level 16 (%d DOF, the same "
                          "per-element work), median of %d after 2 warm-up; the oracle's restatement of "
                          "affine_sshex8_linear_elasticity_apply_macro_adjugate (per macro element: gather, 24x24 "
                          "matrix per micro element, atomic scatter)" % (cb["dofs"], cb["reps"]),
                **{k: cb[k] for k in ("cpu_model", "stream_triad_GBs", "build", "omp", "median_s", "dofs")}}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
