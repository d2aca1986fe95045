#!/usr/bin/env python3
"""Benchmark: sshex8 Laplacian L=8 on a 16^3 macro hex8 unit cube (2,146,689 DOF).

One step = one FP64 matrix-free apply y += A x (x ~ U[-1,1], seed 0) on one MI355X.
Prints ONE JSON line: throughput in DOF/s, the apply kernel's roofline (algorithmic bytes
over its HIP-event-timed average duration), the time to relative residual 1e-10 of plain
FP64 MG-PCG and of FP32/BF16 MG inside FP64 flexible CG (zero Dirichlet on all faces,
RHS ~ U[0,1]), and the CPU baseline: the oracle's OpenMP apply on the host cores.

    python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--no-solve] [--no-cpu]
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


PROFILE = os.path.join(ROOT, "profiles", "sshex8_laplacian_apply_L8_16cube.md")


def profiled_traffic():
    """HBM bytes per apply of this workload measured by rocprofv3 --pmc (FETCH_SIZE and WRITE_SIZE,
    calibrated with a copy kernel of known size), read from the committed profile summary that
    states the command; None without it."""
    try:
        with open(PROFILE) as f:
            m = re.search(r"Total per apply: ([0-9.]+) MB", f.read())
        return float(m.group(1)) * 1e6 if m else None
    except OSError:
        return None


def algorithmic_bytes(N, E, L, value_bytes=8, geom_bytes=4):
    # one read of x, one write of y, int32 connectivity, 6 FFF words per macro element
    return 2 * value_bytes * N + 4 * (L + 1) ** 3 * E + geom_bytes * 6 * E


def cpu_baseline(el, fff, N, L, budget_s=12.0, max_reps=20):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    E = el.shape[1]
    x = np.random.default_rng(0).uniform(-1, 1, N)
    y = np.zeros(N)
    el = np.ascontiguousarray(el)
    for _ in range(2):
        O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, None, x, y)
    times = []
    t_start = time.perf_counter()
    while len(times) < max_reps and (len(times) < 5 or time.perf_counter() - t_start < budget_s):
        t0 = time.perf_counter()
        O.lib.oracle_sshex8_laplacian_apply(L, E, el, fff, None, x, y)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": N / med, "unit": "DOF/s", "cores": int(O.lib.oracle_num_threads()), "kind": "port",
            "sample": f"full apply of this workload ({N} DOF), median of {len(times)} after 2 warm-up; "
                      f"OpenMP restatement of affine_sshex8_laplacian_stencil_apply_fff"}


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
    for name in ("fp64_eager", "fp64", "mixed", "mixed_learned"):
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
    for name in ("fp64", "mixed", "mixed_learned"):
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
    common = dict(mu=mu, lam=lam, macro_box=box, coarse="dense", coeff=coeff, smoother=smoother)
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
        progress("elasticity: training the learned smoother")
        t0 = time.perf_counter()
        data = learned.collect_pairs(h32, n_problems=2, cycles=2, seed=0)
        cnn = S.CNN(16, 4, seed=0)
        learned.train(cnn, data, steps=300, lr=2e-3, batch_elements=16, seed=0)
        hl = gmg.SSGMG(mesh, "elasticity", (adj, det), bc, real_type=S.FLOAT32, learned=cnn, learned_steps=1,
                       **common)
        setups["mixed_learned"] = time.perf_counter() - t0
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
    out["speedup_mixed_vs_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    if with_learned:
        out["speedup_mixed_learned_vs_fp64"] = out["fp64"]["seconds"] / out["mixed_learned"]["seconds"]
    return out


def elasticity_apply_times(S, n_macro, L, steps, coeff_seed=None):
    """Matrix-free sshex8 elasticity apply (y = A x, FP64 and the FP32/BF16 matrix-core operator),
    HIP-event timed over `steps` applies: DOF/s and algorithmic GB/s, B = 2 s_v N + 4 nxe E + 4 * 10 E
    (x read, y written, int32 connectivity, adjugate + det; BF16 geometry: 2 bytes per word)."""
    mel, mpts = S.box_mesh(n_macro + 1, n_macro + 1, n_macro + 1, element_type=S.HEX8)
    mesh, pts_d = S.sshex8_refine_device(L, mel, mpts)
    adj, det = S.mesh_adjugate_det(mesh, pts_d)
    del pts_d
    N, E = 3 * mesh.nnodes, n_macro ** 3
    coeff = None
    if coeff_seed is not None:
        coeff = (10.0 ** np.random.default_rng(coeff_seed).uniform(-1, 1, E)).astype(np.float32)
    nu = 0.3
    mu, lam = 1.0 / (2 * (1 + nu)), nu / ((1 + nu) * (1 - 2 * nu))
    out = {"workload": f"sshex8 linear elasticity apply, L={L} on a {n_macro}^3 macro cube, {N} DOF", "dofs": N,
           "macro_elements": E, "steps": steps}
    for name, rt, dt, vb, gb in (("fp64", S.FLOAT64, np.float64, 8, 4), ("fp32", S.FLOAT32, np.float32, 4, 2)):
        op = S.linear_elasticity(mesh, adj, det, mu, lam, coeff, real_type=rt)
        x = S.DeviceArray.from_numpy(np.random.default_rng(0).uniform(-1, 1, N).astype(dt), dt)
        y = S.DeviceArray(N, dt)
        ms = S.benchmark(op, x, y, warmup=2, reps=steps) / steps
        B = 2 * vb * N + 4 * (L + 1) ** 3 * E + gb * 10 * E
        out[name] = {"ms_per_apply": ms, "dof_per_s": N / (ms * 1e-3), "algorithmic_bytes": B,
                     "GB_per_s": B / (ms * 1e-3) / 1e9, "hbm_frac": B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        del op, x, y
    return out


def elasticity_cases(S, args):
    out = {}
    if args.elasticity_case in ("l8", "both"):
        out["solve_elasticity"] = elasticity_solve_times(S, with_learned=args.elasticity_learned)
    if args.elasticity_case in ("l16", "both"):
        out["solve_elasticity_l16"] = elasticity_solve_times(S, n_macro=24, L=16, coeff_seed=0,
                                                             with_learned=args.elasticity_learned)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-flush", action="store_true")
    ap.add_argument("--no-elasticity", action="store_true", help="skip the ~51M DOF elasticity solve")
    ap.add_argument("--only-elasticity", action="store_true", help="run only the elasticity solve section")
    ap.add_argument("--elasticity-learned", action="store_true",
                    help="also time the learned smoother in the elasticity solve")
    ap.add_argument("--elasticity-case", choices=["l8", "l16", "both"], default="both",
                    help="l8: L=8 on 32^3 macro elements, E=1 (50.9M DOF); l16: L=16 on 24^3 macro elements, "
                         "E log-uniform per macro element (171.2M DOF)")
    ap.add_argument("--mode", choices=["apply", "copy", "elasticity-apply"], default="apply",
                    help="copy: stream-copy 2 GiB of FP64 per step (HBM counter calibration for rocprofv3 --pmc); "
                         "elasticity-apply: time the sshex8 elasticity operators of --elasticity-case")
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
        import time as _t
        S.synchronize()
        t0 = _t.perf_counter()
        for _ in range(args.steps):
            S.lib.sfem_amd_blas_copy(S.FLOAT64, n, S._dptr(a), S._dptr(b), None)
        S.synchronize()
        dt = (_t.perf_counter() - t0) / args.steps
        print(json.dumps({"mode": "copy", "bytes_per_step": 16 * n, "GB/s": 16 * n / dt / 1e9}))
        return
    if args.mode == "elasticity-apply":
        out = {}
        if args.elasticity_case in ("l8", "both"):
            out["elasticity_apply_l8"] = elasticity_apply_times(S, 32, 8, args.steps)
        if args.elasticity_case in ("l16", "both"):
            out["elasticity_apply_l16"] = elasticity_apply_times(S, 24, 16, args.steps, coeff_seed=0)
        print(json.dumps(out))
        return
    if args.only_elasticity:
        print(json.dumps(elasticity_cases(S, args)))
        return
    L, n_macro = 8, 16
    mel, mpts, el, pts, fff = build_problem(n_macro, L)
    N, E = pts.shape[1], el.shape[1]
    mesh = S.Mesh(S.SSHEX8, L, el, N)
    op = S.laplacian(mesh, fff)
    x = S.DeviceArray.from_numpy(np.random.default_rng(0).uniform(-1, 1, N))
    y = S.DeviceArray(N)

    progress(f"apply benchmark: {args.steps} steps")
    total_ms = S.benchmark(op, x, y, warmup=args.warmup, reps=args.steps)
    ms = total_ms / args.steps
    B = algorithmic_bytes(N, E, L)
    achieved = B / (ms * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": N * args.steps / (total_ms * 1e-3),
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
        "config": {"workload": "sshex8 Laplacian apply, level 8 on a 16^3 macro hex8 unit cube "
                               f"({N} DOF, {E} macro elements), FP64 vectors, float32 FFF, int32 ids",
                   "dofs": N, "macro_elements": E, "level": L},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(),
                     "traffic_source": "profiles/sshex8_laplacian_apply_L8_16cube.md (rocprofv3 --pmc FETCH_SIZE, "
                                       "WRITE_SIZE per dispatch, both kernels of one apply)",
                     "algorithmic_bytes_per_apply": B,
                     "note": "working set 46 MB fits the 256 MB MALL: back-to-back applies are cache-hot; "
                             "see roofline_cold for caches overwritten before each apply"},
    }
    if not args.no_flush:
        reps = 20
        cold = S.benchmark(op, x, y, warmup=1, reps=reps, flush_bytes=1 << 30) / reps
        ach = B / (cold * 1e-3) / 1e9
        result["roofline_cold"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": ach / HBM_PEAK_GBS, "ms_per_apply": cold,
                                   "note": "1 GiB written between applies (outside the timed events)"}
    S.set_option("atomic_skeleton", 1)
    try:
        ta = S.benchmark(op, x, y, warmup=args.warmup, reps=args.steps) / args.steps
    finally:
        S.set_option("atomic_skeleton", 0)
    ach = B / (ta * 1e-3) / 1e9
    result["apply_atomic_skeleton"] = {
        "ms_per_apply": ta, "value": N / (ta * 1e-3), "unit": "DOF/s", "achieved": ach, "frac": ach / HBM_PEAK_GBS,
        "note": "option atomic_skeleton: macro-element boundary sums with FP64 atomics inside the line-sweep "
                "kernel (one launch, no partials); not bitwise reproducible, so not the headline value"}
    if not args.no_solve:
        result["solve"] = solve_times(S, mesh, fff, pts, n_macro)
        if not args.no_elasticity:
            result.update(elasticity_cases(S, args))
    if not args.no_cpu:
        progress("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(el, fff, N, L)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
