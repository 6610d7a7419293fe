#!/usr/bin/env python3
"""Benchmark of the per-RK4-stage BEM hot path on one MI355X.

Workload: one sphere, icosphere of frequency nu = 64 (N = 40962 vertices,
M = 81920 triangles), phi = -1 + 0.02 * seeded real spherical harmonics of
degree 1..8, fixed RK4 step dt = 1e-3, free-space (no wall), sigma = 0.

A bench "step" is one RK4 step (four stages, each = assembly of G with the
fused H phi + dense solve + stage update), all state resident in HBM.  The
reported value is ms per RK4 stage of the mixed-precision path (LU whose block
products run on fp16 matrix cores with FP32 accumulation and an FP32 factor,
+ FP64 iterative refinement); the plain FP64 paths (FP64 LU, and FP64
BiCGSTAB, the reference's default solver, at relative residual 1e-10) are
timed on the same state for the speed ratio, and the psi of every path is
compared with the FP64 LU's (target <= 1e-8 relative).

Other workloads (one JSON line each): --workload two-bubbles (two coupled
spheres, N = 20484, mixed vs FP64 over 10 RK4 steps), --workload adjoint (loss
and gradient over 50 RK4 steps at N = 2562), --batch 256 (256 independent
bubbles of uneven size; --batch-nu keeps some size classes).

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bem-bubbles_amd"))
import bemhip  # noqa: E402
import meshgen  # noqa: E402

METRIC = "ms/RK4 stage (assembly+solve) at N=40962; mixed-prec vs FP64 GPU @1e-8 rel"
PEAK = {"lu16_ir64": 2500.0, "lu16x3_ir64": 2500.0 / 3.0, "lu32_ir64": 157.3, "lu64": 78.6}   # dense MFMA TFLOP/s per algorithmic flop
# HBM bytes per timed trailing-update launch at nu = 64 (fp16 LU: the bulk
# two-panel updates; average over the factorisation's launches), from two rocprofv3 --pmc runs (FETCH_SIZE, then
# WRITE_SIZE, each scaled by what it reads for a plain copy of known size):
# profiles/lu16_update_gemm_nu64_hbm_traffic.md (fp16 LU),
# profiles/split16_update_gemm_nu64_counters.md (split-fp16 LU)
TRAFFIC_NU64 = {"lu16_ir64": 6.081e9, "lu16x3_ir64": 6.21e9}
# MFMA busy fraction of the same launches (SQ_VALU_MFMA_BUSY_CYCLES over
# GRBM_GUI_ACTIVE x SIMDs, rocprofv3's MfmaUtil) and the engine clock it implies:
# profiles/lu16_update_gemm_nu64_mfma_utilisation.md, profiles/split16_update_gemm_nu64_counters.md
MFMA_BUSY_NU64 = {"lu16_ir64": (0.329, 1.98), "lu16x3_ir64": (0.517, 1.82)}
DTYPE = {"lu16_ir64": "fp16-mfma lu (fp32 acc, fp32 factor)+f64-ir", "lu16x3_ir64": "f32-lu(split-f16 mfma)+f64-ir",
         "lu32_ir64": "f32-lu+f64-ir", "lu64": "f64", "bicgstab64": "f64"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(threads):
    """The reference's CPU path on the host cores, on a bounded sample:
    assembly (+ H phi) by the oracle (FP64 C/OpenMP restatement of
    ColocSim::assemble_matrices' triangle-parallel loop with its critical
    scatter and 4-pi pass) at nu = 48 (N = 23042), scaled to nu = 64 by the
    work ratio N M; the dense LU by LAPACK dgetrf (scipy, the blocked
    right-looking GEPP that Eigen's PartialPivLU also is) on the leading
    12000 x 12000 block of that G, scaled by (40962 / 12000)^3.  Without
    scipy the oracle's own cache-blocked GEPP takes the LU's place."""
    sys.path.insert(0, ROOT)
    import oracle
    oracle.build()
    v, t = meshgen.icosphere(48)
    n, m = len(v), len(t)
    t0 = time.perf_counter()
    G, H = oracle.assemble(v, t, nthreads=threads)
    t1 = time.perf_counter()
    phi = -1.0 + 0.02 * meshgen.random_sh_field(v, 1, 8, 0)
    oracle.gemv(H, phi, nthreads=threads)
    del H
    t2 = time.perf_counter()
    nl = 12000
    A = np.asfortranarray(G[:nl, :nl])
    del G
    try:
        import scipy.linalg
        lu_kind = "LAPACK dgetrf (scipy)"
        t3 = time.perf_counter()
        scipy.linalg.lu_factor(A, overwrite_a=True, check_finite=False)
    except ImportError:
        lu_kind = "oracle blocked GEPP"
        t3 = time.perf_counter()
        oracle.lu_factor_blocked(A, nthreads=threads, in_place=True)
    t4 = time.perf_counter()
    del A
    N, M = 40962.0, 81920.0
    s_asm = N * M / (n * m)
    s_lu = (N / nl) ** 3
    est_ms = ((t1 - t0) + (t2 - t1)) * s_asm * 1e3 + (t4 - t3) * s_lu * 1e3
    return {"value": round(est_ms, 1), "unit": "ms/stage (N=40962, from a bounded sample)", "cores": threads,
            "kind": "port", "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "sample": f"oracle assembly at nu=48 (N={n}, M={m}) {t1 - t0:.2f}s + H*phi {t2 - t1:.2f}s, scaled by "
                      f"N*M ({s_asm:.2f}x); {lu_kind} of a {nl}x{nl} block of that G {t4 - t3:.2f}s "
                      f"({2.0 / 3.0 * nl ** 3 / (t4 - t3) / 1e9:.0f} GFLOP/s), scaled by (N/{nl})^3 "
                      f"({s_lu:.1f}x); {threads} threads"}


def batch_bench(args):
    """256 independent spheres (nu in {4, 8, 16, 32}, R ~ U[0.5, 2], forcing
    amplitude ~ U[0.1, 0.5] at frequency 0.5, seed 0) advanced together, one
    dense problem per bubble, RK4 dt = 1e-2."""
    cases = meshgen.config_batch(seed=0, count=args.batch)
    if args.batch_nu:
        keep = {int(x) for x in args.batch_nu.split(",")}
        cases = [c for c in cases if int(c.name.rsplit("_nu", 1)[1]) in keep]
    v, t, phi, parts, amp, freq = meshgen.join_batch(cases)
    sizes = np.array([n for _, n in parts], dtype=np.float64)
    N, M = len(v), len(t)
    log(f"[bench] batch of {len(cases)} bubbles: N={N} M={M} solver={args.solver}")
    ctx = bemhip.Context(0)
    sim = bemhip.ColocSim(ctx, v, t, 1.0, 1.0, 0.0, 1.0, solver=args.solver, independent_parts=True)
    sim.set_phi(phi)
    sim.set_part_physics(field_amp=amp, field_freq=freq)
    for w in range(args.warmup):
        t0 = time.perf_counter()
        sim.evolve_system_RK4(1e-2, True)
        log(f"[bench] warmup step {w} {time.perf_counter() - t0:.3f}s")
    tuning = None
    if not args.no_tune:
        # the library's tuner picks the batch schedule on the warm state
        # (outside the timed steps)
        t0 = time.perf_counter()
        tuning = sim.tune()
        kept = [e for e in tuning if e["chosen"]]
        log(f"[bench] tuned in {time.perf_counter() - t0:.1f}s: kept {kept}")
    sim.accum_stats(reset=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.evolve_system_RK4(1e-2, True)
        log(f"[bench] step {k} done {time.perf_counter() - t0:.3f}s")
    wall = time.perf_counter() - t0
    acc = sim.accum_stats(reset=True)
    stages = 4 * args.steps
    ms_stage = wall * 1e3 / stages
    log(f"[bench] batch {ms_stage:.2f} ms/stage {acc}")
    out = {
        "metric": f"ms/RK4 stage (assembly+solve), batch of {len(cases)} independent bubbles",
        "small_problem_solver_max_n": int(os.environ.get("BEMHIP_SMALL_N", "1024")),
        "value": round(ms_stage, 3), "unit": "ms/stage", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.solver],
        "data": "synthetic",
        "config": {"workload": f"{len(cases)} independent spheres, nu in "
                               f"{{{args.batch_nu or '4,8,16,32'}}} (of {args.batch} drawn from {{4,8,16,32}}), R~U[0.5,2], "
                               f"forcing amplitude~U[0.1,0.5] at 0.5, seed 0, RK4 dt=1e-2",
                   "N_total": N, "M_total": M, "solver": args.solver, "stages_timed": stages,
                   "hip_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                   "sum_n2": float((sizes ** 2).sum()), "lu_flops_per_stage": float((2.0 / 3.0 * sizes ** 3).sum())},
        "unknowns_per_s": round(N * stages / wall, 1),
        "phases_ms_per_stage": {k: round(acc[k] / stages, 3) for k in ("ms_assemble", "ms_factor", "ms_solve")},
        "refinement_sweeps_max": acc["iterations"],
    }
    if tuning:
        kept = [e for e in tuning if e["chosen"]][0]
        out["config"]["tuned"] = {"lu_block": kept["nb"], "subclasses": kept["split"],
                                  "lookahead": kept["lookahead"]}
        out["tuning_ms_per_compute_psi"] = {f"nb{e['nb']}_split{e['split']}_la{e['lookahead']}": round(e["ms"], 1)
                                            for e in tuning}
    if not args.no_fp64:
        sim.compute_psi()
        psi_mixed = sim.get_psi()
        sim.set_solver("lu64")
        sim.accum_stats(reset=True)
        t0 = time.perf_counter()
        sim.compute_psi()
        fp64_ms = (time.perf_counter() - t0) * 1e3
        psi64 = sim.get_psi()
        errs = [float(np.max(np.abs(psi_mixed[o:o + n] - psi64[o:o + n])) / np.max(np.abs(psi64[o:o + n])))
                for o, n in parts]
        out["psi_rel_err_vs_fp64_max_over_bubbles"] = max(errs)
        out["fp64_compute_psi_ms"] = round(fp64_ms, 3)
        log(f"[bench] fp64 compute_psi {fp64_ms:.1f} ms, max err {max(errs):.3e}")
    sim.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def two_bubbles_bench(args):
    """Two coupled spheres (icosphere nu = 32 each, N = 20484 unknowns), R = 1.0
    and 0.8 three apart on x, phi = -R * Rdot with Rdot = +0.5 / -0.5, 1e-3
    seeded vertex jitter; 10 RK4 steps of 5e-4 from the same initial state
    with the mixed-precision and the plain FP64 solver, the final states
    compared."""
    c = meshgen.config_two_bubbles(seed=0, nu=32)
    ph = c.physics
    N, M = len(c.verts), len(c.tris)
    ctx = bemhip.Context(0)
    res = {}
    for solver in (args.solver, "lu64"):
        sim = bemhip.ColocSim(ctx, c.verts, c.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma, solver=solver)
        for w in range(args.warmup):   # warmup steps from the initial state, then back to it
            sim.set_phi(c.phi)
            sim.evolve_system_RK4(c.dt, True)
        sim.set_mesh(c.verts, c.tris)
        sim.set_phi(c.phi)
        sim.set_time(0.0)
        sim.accum_stats(reset=True)
        t0 = time.perf_counter()
        for k in range(c.steps):
            sim.evolve_system_RK4(c.dt, True)
        wall = time.perf_counter() - t0
        acc = sim.accum_stats(reset=True)
        x, phi, psi, tm = sim.state()
        res[solver] = (wall * 1e3 / (4 * c.steps), x, phi, psi, acc)
        log(f"[bench] two bubbles {solver}: {res[solver][0]:.2f} ms/stage")
        sim.close()
    ctx.close()
    ms, x, phi, psi, acc = res[args.solver]
    ms64, x64, phi64, psi64, _ = res["lu64"]

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    out = {
        "metric": "ms/RK4 stage (assembly+solve), two coupled bubbles N=20484, mixed vs FP64 over 10 steps",
        "value": round(ms, 3), "unit": "ms/stage", "n_gpus": 1, "steps": c.steps, "warmup": args.warmup,
        "ms_per_step": round(4 * ms, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.solver], "data": "synthetic",
        "config": {"workload": "two spheres nu=32 (N=20484, M=40960), R=1.0 at x=-1.5 and 0.8 at x=+1.5, "
                               "Rdot=+0.5/-0.5, jitter 1e-3 seed 0, RK4 dt=5e-4, 10 steps, coupled (one problem)",
                   "N": N, "M": M, "solver": args.solver,
                   "hip_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "runtime default (4)")},
        "fp64_ms_per_stage": round(ms64, 3), "speedup_vs_fp64": round(ms64 / ms, 3),
        "after_10_steps_rel_diff_vs_fp64": {"x": rel(x, x64), "phi": rel(phi, phi64), "psi": rel(psi, psi64)},
        "refinement_sweeps_per_stage": acc["iterations"] / (4 * c.steps),
    }
    print(json.dumps(out), flush=True)


def adjoint_bench(args):
    """Gradient of the radius-trajectory loss: sphere nu = 16 (N = 2562) at
    rest in p(t) = 1 + 0.3 sin(pi t), 50 RK4 steps of 1e-2, checkpoints every
    10 steps; the target trajectory from the run whose initial phi carries a
    seeded 1 % spherical-harmonic perturbation.  Times one loss + gradient
    (w.r.t. x0 and phi0) per solver against the 50 forward RK4 steps alone."""
    c = meshgen.config_adjoint(seed=0)
    ph = c.physics
    ctx = bemhip.Context(0)
    out_s = {}
    fwd = bemhip.ColocSim(ctx, c.verts, c.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma, field_amp=ph.field_amp,
                          field_freq=ph.field_freq, solver="lu64")
    V0 = fwd.get_volume()
    fwd.set_phi(meshgen.adjoint_target_phi(c, seed=1))
    target = []
    for _ in range(c.steps):
        fwd.evolve_system_RK4(c.dt, True)
        target.append((3.0 * fwd.get_volume() / (4.0 * np.pi)) ** (1.0 / 3.0))
    target = np.array(target)
    fwd.close()
    for solver in (args.solver, "lu64"):
        sim = bemhip.ColocSim(ctx, c.verts, c.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma, field_amp=ph.field_amp,
                              field_freq=ph.field_freq, solver=solver, V_0=V0)
        sim.set_phi(c.phi)
        sim.adjoint_radius(c.steps, c.dt, target, every=10)   # warmup
        t0 = time.perf_counter()
        L, R, gx, gp = sim.adjoint_radius(c.steps, c.dt, target, every=10)
        t_grad = time.perf_counter() - t0
        # the forward alone: the same 50 RK4 steps from the same state
        sim.set_time(0.0)
        t0 = time.perf_counter()
        for _ in range(c.steps):
            sim.evolve_system_RK4(c.dt, True)
        t_fwd = time.perf_counter() - t0
        out_s[solver] = {"grad_ms": round(t_grad * 1e3, 1), "forward_ms": round(t_fwd * 1e3, 1),
                         "grad_over_forward": round(t_grad / t_fwd, 2), "loss": L}
        log(f"[bench] adjoint {solver}: {out_s[solver]}")
        sim.close()
    ctx.close()
    m = out_s[args.solver]
    out = {
        "metric": "ms per loss+gradient (50 RK4 steps, checkpoints every 10), sphere N=2562",
        "value": m["grad_ms"], "unit": "ms/gradient", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": m["grad_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 state, " + DTYPE[args.solver] + " solves", "data": "synthetic",
        "config": {"workload": "sphere nu=16 (N=2562) at rest, p_inf(t)=1+0.3 sin(pi t), 50 RK4 steps dt=1e-2, "
                               "checkpoint every 10, target from 1% perturbed phi (seed 1)",
                   "hip_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "runtime default (4)")},
        "per_solver": out_s,
    }
    print(json.dumps(out), flush=True)


def exterior_bench(args):
    """Potential off the surface (compute_exterior_pot): the nu = 64 sphere
    (M = 81920 panels) with its solved psi, at 65536 targets on a 256 x 256
    grid of the plane z = 0.3 over [-3, 3]^2; 19-point rule per (target,
    panel).  Times bemhip_exterior_potential on the solved psi (inputs
    uploaded per call: 2.6 MB)."""
    case = meshgen.config_large(seed=0, nu=64)
    ph = case.physics
    ctx = bemhip.Context(0)
    sim = bemhip.ColocSim(ctx, case.verts, case.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma, solver=args.solver)
    sim.set_phi(case.phi)
    sim.compute_psi()
    psi = sim.get_psi()
    sim.close()
    g = np.linspace(-3.0, 3.0, 256)
    X, Y = np.meshgrid(g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 0.3)], axis=1)
    # the standalone entry point (compute_exterior_pot) on the solved psi;
    # the simulation's exterior_pot would solve for psi again first
    for _ in range(max(1, args.warmup)):
        ctx.exterior_potential(case.verts, case.tris, case.phi, psi, pts)
    reps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(reps):
        u = ctx.exterior_potential(case.verts, case.tris, case.phi, psi, pts)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    ctx.close()
    P, M = pts.shape[0], len(case.tris)
    evals = 19.0 * P * M
    flops = 14.0 * evals
    out = {
        "metric": "ms per exterior-potential evaluation (65536 targets x 81920 panels, 19-point rule)",
        "value": round(ms, 3), "unit": "ms/evaluation", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "sphere nu=64 (M=81920) with solved psi, 256x256 targets on z=0.3 over [-3,3]^2",
                   "targets": P, "panels": M, "hip_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")},
        "kernel_evaluations_per_s": evals / (ms * 1e-3),
        "roofline": {"bound": "fp64 valu", "achieved": round(flops / (ms * 1e-3) / 1e12, 2), "peak": 78.6,
                     "unit": "TFLOP/s", "frac": round(flops / (ms * 1e-3) / 78.6e12, 4),
                     "note": "14 FP64 ops per kernel evaluation (incl. one rsqrt); wall time incl. the 2.6 MB of inputs uploaded per call"},
        "finite": bool(np.all(np.isfinite(u))),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nu", type=int, default=64)
    ap.add_argument("--solver", default="lu16_ir64", choices=sorted(bemhip.SOLVERS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="batch: keep the default schedule (no tuner run)")
    ap.add_argument("--workload", default="sphere", choices=["sphere", "two-bubbles", "adjoint", "exterior"],
                    help="sphere: the N=40962 RK4 stage (default); two-bubbles: N=20484 coupled pair, "
                         "mixed vs FP64 over 10 steps; adjoint: loss + gradient over 50 steps at N=2562; "
                         "exterior: potential at 65536 points off the N=40962 sphere")
    ap.add_argument("--batch", type=int, default=0,
                    help="instead of the single-sphere workload, time a batch of this many independent bubbles")
    ap.add_argument("--batch-nu", default="",
                    help="keep only the batch's bubbles of these icosphere frequencies (comma list, e.g. 4,8)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); multi-GPU sharding is "
              f"design-only in this build", file=sys.stderr)
        sys.exit(2)
    if args.batch:
        batch_bench(args)
        return
    if args.workload == "two-bubbles":
        two_bubbles_bench(args)
        return
    if args.workload == "adjoint":
        adjoint_bench(args)
        return
    if args.workload == "exterior":
        exterior_bench(args)
        return

    case = meshgen.config_large(seed=0, nu=args.nu)
    N, M = len(case.verts), len(case.tris)
    log(f"[bench] nu={args.nu} N={N} M={M} solver={args.solver}")
    ctx = bemhip.Context(0)
    ph = case.physics
    sim = bemhip.ColocSim(ctx, case.verts, case.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma,
                          solver=args.solver)
    sim.set_phi(case.phi)

    for w in range(args.warmup):
        t0 = time.perf_counter()
        sim.evolve_system_RK4(case.dt, True)
        log(f"[bench] warmup step {w} {time.perf_counter() - t0:.3f}s")
    sim.accum_stats(reset=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.evolve_system_RK4(case.dt, True)
        log(f"[bench] step {k} done {time.perf_counter() - t0:.3f}s")
    wall = time.perf_counter() - t0
    acc = sim.accum_stats(reset=True)
    stages = 4 * args.steps
    ms_stage = wall * 1e3 / stages
    log(f"[bench] mixed {ms_stage:.2f} ms/stage {acc}")

    out = {
        "metric": METRIC, "value": round(ms_stage, 3), "unit": "ms/stage", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.solver],
        "data": "synthetic",
        "config": {"workload": f"one sphere, nu={args.nu} icosphere (N={N}, M={M}), phi=-1+0.02*SH(l=1..8) seed 0, "
                               f"RK4 dt={case.dt}, one bench step = one RK4 step = 4 stages",
                   "N": N, "M": M, "solver": args.solver, "stages_timed": stages,
                   "hip_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "runtime default (4)")},
        "phases_ms_per_stage": {k: round(acc[k] / stages, 3) for k in ("ms_assemble", "ms_factor", "ms_solve")},
        "refinement_sweeps_per_stage": acc["iterations"] / stages,
        "lu_panels": acc["panels"], "lu_panels_pivoted": acc["panels_pivoted"],
    }
    gf, gms = acc["gemm_flops"], acc["gemm_ms"]
    if gms > 0:
        achieved = gf / (gms * 1e-3) / 1e12
        launches = acc["gemm_launches"]
        merged = os.environ.get("BEMHIP_LU16_MERGE", "1") != "0"
        rl = {"bound": "mfma", "kernel": "LU trailing update " + (
                  ("bemhip::k_mm16<0,0,true> (bulk update of two panels, K = 2048" if merged else
                   "bemhip::k_mm16<0,0> (K = 1024") + ", fp16 MFMA, FP32 accumulation)"
                  if args.solver == "lu16_ir64" else "(rocBLAS GEMM)"),
              "achieved": round(achieved, 2), "peak": round(PEAK[args.solver], 1), "unit": "TFLOP/s",
              "frac": round(achieved / PEAK[args.solver], 4),
              "traffic": TRAFFIC_NU64.get(args.solver) if args.nu == 64 else None,
              "launches": launches, "avg_launch_ms": round(gms / max(1, launches), 4)}
        if args.solver == "lu16_ir64":
            # algorithmic bytes: C read and written in FP32, the fp16 L21 and U12
            # tiles read once, over the factorisation's update launches
            nb = int(os.environ.get("BEMHIP_NB16", "1024"))
            shapes = lu_update_shapes(N, nb, os.environ.get("BEMHIP_LU16_MERGE", "1") != "0",
                                      int(os.environ.get("BEMHIP_LU16_BULK_FIRST", "25000")))
            ab = sum(8.0 * m * n + 2.0 * (m + n) * k for m, n, k in shapes)
            fl = sum(2.0 * m * n * k for m, n, k in shapes)
            if abs(fl - gf / stages) < 1e-6 * fl:
                rl["algorithmic_bytes_per_launch"] = round(ab / len(shapes))
                rl["hbm_achieved_GBps"] = round(ab * stages / (gms * 1e-3) / 1e9, 1)
                rl["hbm_frac"] = round(ab * stages / (gms * 1e-3) / 8e12, 4)
        if args.nu == 64 and args.solver in MFMA_BUSY_NU64:
            busy, clk = MFMA_BUSY_NU64[args.solver]
            rl["mfma_busy_pmc"] = busy
            rl["engine_clock_ghz_pmc"] = clk
        out["roofline"] = rl

    # assembly: algorithmic output bytes (G in FP64, plus its FP32 copy for the
    # mixed solvers) and algorithmic flops (35 per quadrature point, 7 points
    # per collocation-row x triangle pair) over its device time
    asm_ms = acc["ms_assemble"] / stages
    g_bytes = (8.0 + (4.0 if args.solver != "lu64" else 0.0)) * N * N
    asm_flops = 35.0 * 7.0 * N * M
    out["assembly"] = {"kernels": "k_entry_prep + k_far + k_self", "ms_per_stage": round(asm_ms, 3),
                       "achieved_GBps": round(g_bytes / (asm_ms * 1e-3) / 1e9, 1),
                       "hbm_frac": round(g_bytes / (asm_ms * 1e-3) / 8e12, 4),
                       "achieved_fp64_TFLOPs": round(asm_flops / (asm_ms * 1e-3) / 1e12, 2),
                       "fp64_frac": round(asm_flops / (asm_ms * 1e-3) / 78.6e12, 4), "bound": "fp64 valu"}

    # the plain FP64 paths on the same state: the FP64 LU (the ratio of the
    # metric) and BiCGSTAB (the reference's default solver) at relative
    # residual 1e-10; psi of every path against the FP64 LU's
    if not args.no_fp64:
        sim.compute_psi()
        psi_mixed = sim.get_psi()
        sim.set_solver("lu64")
        sim.compute_psi()
        psi64 = sim.get_psi()
        err = float(np.max(np.abs(psi_mixed - psi64)) / np.max(np.abs(psi64)))
        sim.set_solver("bicgstab64", tol=1e-10)
        sim.compute_psi()
        errb = float(np.max(np.abs(sim.get_psi() - psi64)) / np.max(np.abs(psi64)))
        # the FP32-class mixed solver (every block product as three fp16
        # products, FP32 accumulation) on the same state
        x3 = "lu16x3_ir64"
        if args.solver != x3:
            sim.set_solver(x3)
            sim.compute_psi()
            errx3 = float(np.max(np.abs(sim.get_psi() - psi64)) / np.max(np.abs(psi64)))
        sim.set_solver("lu64")
        sim.accum_stats(reset=True)
        t0 = time.perf_counter()
        sim.evolve_system_RK4(case.dt, True)
        fp64_ms = (time.perf_counter() - t0) * 1e3 / 4
        a64 = sim.accum_stats(reset=True)
        log(f"[bench] fp64 lu {fp64_ms:.2f} ms/stage err={err:.3e} {a64}")
        out["fp64_ms_per_stage"] = round(fp64_ms, 3)
        out["speedup_vs_fp64"] = round(fp64_ms / ms_stage, 3)
        out["psi_rel_err_vs_fp64"] = err
        if a64["gemm_ms"] > 0:
            out["fp64_gemm_tflops"] = round(a64["gemm_flops"] / (a64["gemm_ms"] * 1e-3) / 1e12, 2)
        sim.set_solver("bicgstab64", tol=1e-10)
        sim.accum_stats(reset=True)
        t0 = time.perf_counter()
        sim.evolve_system_RK4(case.dt, True)
        bcg_ms = (time.perf_counter() - t0) * 1e3 / 4
        ab = sim.accum_stats(reset=True)
        log(f"[bench] fp64 bicgstab {bcg_ms:.2f} ms/stage err={errb:.3e} {ab}")
        out["fp64_bicgstab"] = {"ms_per_stage": round(bcg_ms, 3), "tol": 1e-10,
                                "iterations_per_stage": ab["iterations"] / 4.0,
                                "psi_rel_err_vs_fp64_lu": errb,
                                "speedup_of_mixed": round(bcg_ms / ms_stage, 3)}
        if args.solver != x3:
            sim.set_solver(x3)
            sim.accum_stats(reset=True)
            t0 = time.perf_counter()
            sim.evolve_system_RK4(case.dt, True)
            x3_ms = (time.perf_counter() - t0) * 1e3 / 4
            ax3 = sim.accum_stats(reset=True)
            log(f"[bench] fp32-class {x3} {x3_ms:.2f} ms/stage err={errx3:.3e} {ax3}")
            out["fp32_class_leg"] = {"solver": x3, "dtype": DTYPE[x3], "ms_per_stage": round(x3_ms, 3),
                                     "psi_rel_err_vs_fp64_lu": errx3,
                                     "speedup_vs_fp64": round(fp64_ms / x3_ms, 3),
                                     "refinement_sweeps_per_stage": ax3["iterations"] / 4.0}
    sim.close()
    ctx.close()

    if not args.no_cpu:
        try:
            threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
        except (AttributeError, ValueError):
            threads = os.cpu_count() or 1
        log(f"[bench] cpu baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(out), flush=True)


def lu_update_shapes(N, nb, merged=True, bulk_first=25000):
    """(m, n, k) of the fp16 LU's timed update launches.  Merged (the library's
    default): per pair of panels (k, k+1) the bulk update of every column
    beyond panel k+3 by both panels (K = 2 nb); while at least `bulk_first`
    rows are left after panel k+1 in two launches, the rows of panels k+2 and
    k+3 first, then the rows beyond, later in one (the narrow products are not
    timed); step by step: per panel the next panel's columns, then the rest."""
    out = []
    k = 0
    w = lambda r: min(nb, r)
    if not merged:
        while k < N:
            jb = w(N - k)
            rest = N - k - jb
            if rest <= 0:
                break
            jb2 = w(rest)
            out.append((rest, jb2, jb))
            if rest > jb2:
                out.append((rest, rest - jb2, jb))
            k += jb
        return out
    jb = w(N)
    while True:
        rest = N - k - jb
        if rest <= 0:
            break
        jb2 = w(rest)
        rest2 = rest - jb2
        if rest2 <= 0:
            break
        jb3 = w(rest2)
        rest3 = rest2 - jb3
        if rest3 > 0:
            jb4 = w(rest3)
            rest4 = rest3 - jb4
            if rest4 > 0:
                if rest2 >= bulk_first:
                    out.append((jb3 + jb4, rest4, jb + jb2))
                    out.append((rest4, rest4, jb + jb2))
                else:
                    out.append((rest2, rest4, jb + jb2))
        k += jb + jb2
        jb = jb3
    return out


if __name__ == "__main__":
    main()
