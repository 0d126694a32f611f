#!/usr/bin/env python3
"""Benchmark: time-to-converge and LM iter/s of the MI355X frame-BA backend.

Workload: aerial block, 1000 cameras in 3 strips at 300 m, 200k points on an fBm terrain, ~1.5M observations
(seed 2), Cauchy 1 px, shared focal + Brown distortion refined over 4 groups, online point Schur. A "step" is one
complete solve from the initial state to the reference's convergence stop (step RMS <= 1e-6,
BundleAdjustPlaMatrix.cpp:446-466,524-526; at most 400 LM iterations) with the inputs resident in HBM (the problem
structure is uploaded once, before the timed region).

Reported: the mixed-precision path (FP32 Jacobians/Schur, FP64 cost and acceptance; FP64 assembly once FP32 steps
reach their rounding floor) as `value` (LM iter/s) and `ms_per_step` (time-to-converge), the plain FP64 GPU path
beside it (both with the dense reduced solve), the time each needs to bring the objective within 1e-6 of the
converged FP64 oracle's, the reference's GPU algorithm (FP64 block-Jacobi PCG to 2e-10) and the truncated-Newton
FP32 PCG configuration over the reference's 20-iteration budget, the relative final-cost error and the refined
states against the converged oracle (tests/golden), the roofline of the dominant kernel family (HIP events inside
the library, on its stream), the CPU oracle timed on the same workload on all host cores (and a one-thread sample),
and the batched-window, adjoint and quality legs.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
FAMILIES = ("linearize_points", "camera_direct", "schur_blocks", "pcg_iteration", "cost", "back_substitution",
            "dense_factor", "dense_solve", "band_factor", "band_solve")
MFMA_FAMILIES = ("dense_factor", "dense_solve", "band_factor", "band_solve")
METRIC = "time-to-converge + LM iter/s, 1.5M obs, mixed vs FP64 GPU at <=1e-6 rel cost"


def fingerprint(p) -> str:
    import numpy as np

    h = hashlib.sha256()
    for a in (p.cam_rotation, p.cam_center, p.cam_intrinsics, p.track_point, p.track_offset, p.obs_camera, p.obs_uv):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(family: str, problem_fingerprint: str):
    """HBM bytes per launch of a kernel family from the committed rocprofv3 counter summaries in profiles/
    (lines "traffic_bytes_per_launch[<family>]: N" and "workload_fingerprint: F"), or None."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith((".md", ".txt")):
            continue
        text = open(os.path.join(pdir, name)).read()
        if f"workload_fingerprint: {problem_fingerprint}" not in text:
            continue
        for line in text.splitlines():
            key = f"traffic_bytes_per_launch[{family}]:"
            if line.strip().startswith(key):
                return float(line.split(":", 1)[1])
    return None


def copy_calibration(nbytes: int) -> None:
    """Three device-to-device copies of nbytes through the HIP runtime: a kernel of known traffic
    (nbytes read, nbytes written) that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for buf in (src, dst):
        if hip.hipMalloc(C.byref(buf), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    for _ in range(3):
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)


def batch_leg(num_windows: int, precisions, oracle_sample: int, warmup: bool = True) -> dict:
    """Batched local windows through pb_hip_solve_batch: 512 windows of U{8..64} cameras, U{500..8000} points,
    track lengths U{2..12} (seed 4). Reports the batch solve time, the per-window device time to converge
    (p50/p99) and the final-cost error of a sample of windows against the FP64 oracle."""
    import numpy as np

    from plabundle_amd import options as PO
    from plabundle_amd import synthetic as S
    from plabundle_amd.solver import Solver, SolverWorkspace

    t = time.time()
    wins = S.batch_windows(seed=4, num_windows=num_windows)
    nobs = int(sum(w.num_observations for w in wins))
    log(f"batch: generated {num_windows} windows, {nobs} observations in {time.time() - t:.1f}s")
    o = S.window_options(PO.SolveOptions(), max_iterations=20)
    ws = SolverWorkspace(0)
    out = {"workload": f"{num_windows} local windows: cameras U{{8..64}}, points U{{500..8000}}, track length "
                       "U{2..12}, one shared focal + Brown group per window, Huber 1 px, at most 20 LM iterations, "
                       "seed 4",
           "windows": num_windows, "observations": nobs,
           "linear_solve": "dense FP64 Cholesky of each window's reduced camera system (one workgroup per window)"}
    sample = list(np.random.default_rng(4).choice(num_windows, size=min(oracle_sample, num_windows), replace=False))
    refs = {}
    if sample:
        from oracle import oracle as O

        co = PO.to_c(o)
        for w in sample:
            refs[w] = O.Oracle(wins[w], threads=O.lib().oracle_max_threads()).solve(co)
    if warmup:
        # one untimed batch first: it times the launch shapes of the window factorisation for both assembly
        # precisions, which the timed batches (same batch class) then reuse
        _, summ = Solver().solve_batch(wins, o, ws, precision=PO.Precision.Mixed)
        out["warmup_batch_solve_s"] = round(summ["solve_seconds"], 4)
        log(f"batch warm-up (times the factorisation shapes): {out['warmup_batch_solve_s']} s")
    for prec in precisions:
        name = "mixed" if prec == PO.Precision.Mixed else "fp64"
        res, summ = Solver().solve_batch(wins, o, ws, precision=prec)
        if summ["abi_status"] != 0:
            out[name] = {"error": summ["message"]}
            continue
        dev = np.array([r.report["device_seconds"] for r in res])
        its = np.array([r.report["iterations"] for r in res])
        work = float(sum(r.report["observation_count"] * r.report["iterations"] for r in res))
        errs = [abs(res[w].finalCost - refs[w]["final_cost"]) / refs[w]["final_cost"] for w in sample]
        out[name] = {
            "solve_s": round(summ["solve_seconds"], 4),
            "setup_s": round(summ["setup_seconds"], 2),
            "rounds": summ["iterations"],
            "converged_windows": int(sum(r.report["converged"] for r in res)),
            "window_time_to_converge_p50_ms": round(float(np.percentile(dev, 50)) * 1e3, 3),
            "window_time_to_converge_p99_ms": round(float(np.percentile(dev, 99)) * 1e3, 3),
            "window_time_to_converge_max_ms": round(float(dev.max()) * 1e3, 3),
            "window_lm_iterations_mean": round(float(its.mean()), 3),
            "window_fp64_iterations_mean": round(float(np.mean([r.report["fp64_iterations"] for r in res])), 3),
            "windows_per_s": round(num_windows / summ["solve_seconds"], 2),
            "observation_iterations_per_s": round(work / summ["solve_seconds"], 1),
            "factor_tuning": summ["factor_tuning"],
            "oracle_sample_max_rel_cost_err": float(max(errs)) if errs else None,
            "oracle_sample": [{"window": int(w), "rel_cost_err": float(e), "iterations": res[w].report["iterations"],
                               "fp64_iterations": res[w].report["fp64_iterations"],
                               "status": int(res[w].report["solve_status"]),
                               "oracle_iterations": refs[w]["iterations"], "oracle_status": refs[w]["solve_status"]}
                              for w, e in zip(sample, errs)],
        }
        log(f"batch {name}: {out[name]}")
    return out


def survey_block_leg(strips: int, cameras_per_strip: int, num_points: int, steps: int):
    """The large aerial survey block (20 strips x 500 cameras, full Metashape calibration over 8 groups, 2M points,
    about 15M observations with 2 % outliers, 50 control points, 10k LiDAR planes; path A): 20 LM iterations mixed,
    and FP64 with the same PCG stopping rule as the precision reference. No CPU oracle at this size; the control
    points' residuals before and after are reported as a check of the datum."""
    import numpy as np

    from plabundle_amd import options as PO
    from plabundle_amd import synthetic as S
    from plabundle_amd.solver import Solver, SolverWorkspace

    t = time.time()
    problem = S.large_aerial_problem(strips=strips, cameras_per_strip=cameras_per_strip, num_points=num_points)
    gen_s = time.time() - t
    log(f"survey block: {problem.num_cameras} cameras, {problem.num_tracks} tracks, {problem.num_observations} "
        f"observations, {problem.num_control_points} control points, {problem.num_laser_planes} laser planes "
        f"generated in {gen_s:.1f}s")
    ws = SolverWorkspace(0)
    solver = Solver()

    def run(precision, inexact=False, profile=False):
        o = S.large_aerial_options(PO.SolveOptions())
        o.backend.profileKernels = profile
        o.backend.mixedSwitchDecrease = -1.0  # the truncated-Newton FP32 PCG configuration throughout
        if inexact:
            o.backend.pcgRelativeTolerance = 1e-3
            o.backend.pcgModelEta = 0.1
            o.backend.pcgMaxIterations = max(200, 3 * (problem.num_cameras + 24) // 5)
        r = solver.solve(problem, o, ws, precision=precision)
        if not r.usable():
            raise RuntimeError(f"survey block solve failed: {r.status.name} {r.backendMessage}")
        return r

    def cp_rms(X):
        d = X[problem.cp_track] - problem.cp_point
        return float(np.sqrt(np.mean(np.sum(d * d, axis=1))))

    first = run(PO.Precision.Mixed)
    log(f"survey block warmup: {first.report['iterations']} it, setup {first.report['setup_seconds']:.2f}s, "
        f"solve {first.report['solve_seconds']:.3f}s, cost {first.finalCost:.10g}")
    mixed = []
    for i in range(steps):
        r = run(PO.Precision.Mixed, profile=(i == 0))
        mixed.append(r)
        log(f"survey block mixed {i}: {r.report['iterations']} it, {r.report['linear_iterations']} PCG it, "
            f"solve {r.report['solve_seconds']:.3f}s, cost {r.finalCost:.10g}")
    fams = {}
    for fam in ("linearize_points", "camera_direct", "schur_blocks", "pcg_iteration", "cost", "back_substitution"):
        st = ws.kernel_stats(fam)
        if st["launches"]:
            avg = st["total_ms"] / st["launches"]
            fams[fam] = {"timed_launches": st["launches"], "avg_ms": round(avg, 5),
                         "GBps": round(st["bytes_per_launch"] / (avg * 1e-3) / 1e9, 1) if avg > 0 else 0.0}
    ref = run(PO.Precision.Fp64, inexact=True)
    log(f"survey block fp64 (same PCG stop): {ref.report['iterations']} it, "
        f"{ref.report['linear_iterations']} PCG it, solve {ref.report['solve_seconds']:.3f}s, cost {ref.finalCost:.10g}")
    mt = sum(r.report["solve_seconds"] for r in mixed) / len(mixed)
    out = {
        "workload": f"survey block: {strips} strips x {cameras_per_strip} cameras at 300 m, full Metashape calibration "
                    f"over 8 groups, {num_points} points, 2% outliers, 50 control points, 10k LiDAR planes, seed 3",
        "observations": problem.num_observations,
        "generation_s": round(gen_s, 1),
        "setup_s": round(first.report["setup_seconds"], 3),
        "path": "assembled damped Schur (laser planes)" if not mixed[0].report["online_schur_used"] else "online Schur",
        "mixed_time_s": round(mt, 4),
        "mixed_lm_iter_per_s": round(mixed[0].report["iterations"] / mt, 3),
        "mixed_pcg_iterations": mixed[0].report["linear_iterations"],
        "mixed_message": mixed[0].report["message"],
        "fp64_same_stop_time_s": round(ref.report["solve_seconds"], 4),
        "fp64_same_stop_pcg_iterations": ref.report["linear_iterations"],
        "initial_cost": mixed[0].initialCost,
        "final_cost_mixed": mixed[0].finalCost,
        "final_cost_fp64_same_stop": ref.finalCost,
        "rel_cost_diff_mixed_vs_fp64": abs(mixed[0].finalCost - ref.finalCost) / ref.finalCost,
        "control_rms_m": {"initial": round(cp_rms(problem.track_point), 5), "mixed": round(cp_rms(mixed[0].track_point), 5),
                          "fp64": round(cp_rms(ref.track_point), 5)},
        "families": fams,
    }
    ws.close()
    return out


def state_differences(ref, r):
    """Metrics of comparePhotoAlignmentOutcomes (src/photo_alignment.cpp:225-318) against the converged golden
    state: largest centre distance, largest rotation angle (of R_ref^T R), largest point distance over the sampled
    tracks valid on both sides, mean height difference, and sampled valid-flag mismatches."""
    import numpy as np

    tr = np.einsum("ij,ij->i", ref["cam_rotation"], r.cam_rotation)
    rot = np.degrees(np.arccos(np.clip(0.5 * (tr - 1.0), -1.0, 1.0)))
    idx = ref["point_index"]
    va, vb = ref["track_valid"] != 0, r.track_valid[idx] != 0
    both = va & vb
    pd = np.linalg.norm(ref["track_point"][both] - r.track_point[idx][both], axis=1)
    gi = ref["group_intrinsics"]
    f = r.cam_intrinsics[ref["group_cameras"], 0]
    return {"max_camera_center_difference_m": float(np.linalg.norm(ref["cam_center"] - r.cam_center, axis=1).max()),
            "mean_camera_height_difference_m": float(np.mean(r.cam_center[:, 2] - ref["cam_center"][:, 2])),
            "max_camera_rotation_difference_deg": float(rot.max()),
            "max_point_difference_m_sampled": float(pd.max()) if pd.size else 0.0,
            "max_focal_relative_difference": float(np.max(np.abs(f - gi[:, 0]) / np.abs(gi[:, 0]))),
            "valid_track_mask_mismatches_sampled": int(np.count_nonzero(va != vb))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed converged mixed solves")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--max-iterations", type=int, default=400,
                    help="LM budget of the converged solves (the reference's stop ends them long before)")
    ap.add_argument("--pcg-iterations", type=int, default=20,
                    help="LM iterations of the FP64 PCG and truncated-Newton legs (0 = skip)")
    ap.add_argument("--cpu-iterations", type=int, default=20,
                    help="LM iterations of the CPU oracle run on all host cores")
    ap.add_argument("--cpu-single-iterations", type=int, default=1,
                    help="LM iterations of the one-thread CPU oracle sample (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--small", action="store_true", help="reduced problem for quick checks (not a bench line)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint and quality timing")
    ap.add_argument("--batch-windows", type=int, default=512,
                    help="windows of the batched local-BA leg (0 = skip)")
    ap.add_argument("--batch-fp64", type=int, default=1, help="also solve the batch in FP64")
    ap.add_argument("--batch-oracle-sample", type=int, default=8,
                    help="windows of the batch checked against the FP64 oracle")
    ap.add_argument("--batch-warmup", type=int, default=1,
                    help="solve the batch once untimed first (times the window factorisation shapes)")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched-window leg and print its JSON (profiling the batch)")
    ap.add_argument("--survey-block", action="store_true",
                    help="run only the large survey-block leg (15M observations, path A) and print its JSON")
    ap.add_argument("--survey-strips", type=int, default=20)
    ap.add_argument("--survey-cameras", type=int, default=500, help="cameras per strip of the survey block")
    ap.add_argument("--survey-points", type=int, default=2_000_000)
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only copy this many MiB device-to-device three times (the counter calibration run)")
    ap.add_argument("--profile-only", choices=("mixed", "fp64"), default=None,
                    help="only the warm-up and one converged solve of this path (profiler runs)")
    args = ap.parse_args()
    if args.copy_calibration_mb > 0:
        copy_calibration(args.copy_calibration_mb << 20)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this backend runs on one MI355X; "
              "the multi-GPU sharding plan is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from plabundle_amd import options as PO
    from plabundle_amd import synthetic as S
    from plabundle_amd.solver import Solver, SolverWorkspace

    if args.survey_block:
        print(json.dumps({"survey_block": survey_block_leg(args.survey_strips, args.survey_cameras,
                                                            args.survey_points, args.steps)}), flush=True)
        return

    if args.batch_only:
        precs = [PO.Precision.Mixed] + ([PO.Precision.Fp64] if args.batch_fp64 else [])
        print(json.dumps({"batch": batch_leg(args.batch_windows or 512, precs, args.batch_oracle_sample,
                                                   bool(args.batch_warmup))}), flush=True)
        return

    t = time.time()
    if args.small:
        problem = S.aerial_problem(cameras_per_strip=(60, 60, 60), num_points=30000)
    else:
        problem = S.aerial_problem()
    log(f"generated {problem.num_cameras} cameras, {problem.num_tracks} tracks, {problem.num_observations} "
        f"observations in {time.time() - t:.1f}s")
    ws = SolverWorkspace(0)
    solver = Solver()

    def solve(precision, max_iterations=args.max_iterations, route=PO.ReducedSolver.Auto, profile=False,
              trace=False, configure=None):
        o = S.aerial_options(PO.SolveOptions())
        o.solver.maxIterations = max_iterations
        o.backend.reducedSolver = route
        o.backend.profileKernels = profile
        if configure:
            configure(o)
        points = []
        if trace:
            def progress(sm):
                points.append((sm.currentIteration, sm.averageRmsPixels, time.perf_counter()))
                return True

            o.solver.progressCallback = progress
        t0 = time.perf_counter()
        r = solver.solve(problem, o, ws, precision=precision)
        if not r.usable():
            raise RuntimeError(f"solve failed: {r.status.name} {r.backendMessage}")
        return r, [(it, rms * rms * r.report["observation_count"], tt - t0) for it, rms, tt in points]

    def brief(r):
        rp = r.report
        return (f"{rp['iterations']} it ({rp['accepted_steps']} accepted), solve {rp['solve_seconds']:.4f}s, "
                f"cost {r.finalCost:.12g}, {rp['linear_solver_name']}, status {r.status.name}")

    for i in range(args.warmup):
        for prec in (PO.Precision.Mixed, PO.Precision.Fp64):
            r, _ = solve(prec, max_iterations=2)
            log(f"warmup {i} {prec.name}: {brief(r)} (setup {r.report['setup_seconds']:.2f}s)")

    if args.profile_only:
        prec = PO.Precision.Mixed if args.profile_only == "mixed" else PO.Precision.Fp64
        r, _ = solve(prec)
        log(f"profile-only {prec.name}: {brief(r)}")
        print(json.dumps({"profile_only": args.profile_only, "iterations": r.report["iterations"],
                          "solve_seconds": r.report["solve_seconds"]}), flush=True)
        return

    # timed converged solves: the mixed path (the headline), then the FP64 path
    mixed = []
    for i in range(args.steps):
        r, _ = solve(PO.Precision.Mixed, profile=(i == 0))
        mixed.append(r)
        log(f"step {i}: mixed {brief(r)}")
    stats = {fam: ws.kernel_stats(fam) for fam in FAMILIES}
    # the FP64 path timed the same way (the first of the steps carries the kernel-family events)
    fp64s = []
    for i in range(args.steps):
        r, _ = solve(PO.Precision.Fp64, profile=(i == 0))
        fp64s.append(r)
        if i == 0:
            stats64 = {fam: ws.kernel_stats(fam) for fam in FAMILIES}
        log(f"step {i}: fp64 {brief(r)}")
    fp64 = fp64s[-1]
    seed, _ = solve(PO.Precision.Fp32Seed)
    log(f"fp32-seed (reference mixed mode): {brief(seed)}; seeds accepted {seed.report['mixed_seed_steps']}, "
        f"rejected {seed.report['dense_fallbacks']}")

    # time to bring the objective within 1e-6 of the converged oracle's (traced runs: a host callback per step)
    golden = gstate = None
    gpath = os.path.join(ROOT, "tests", "golden", "aerial_oracle.json")
    if not args.small and os.path.exists(gpath):
        golden = json.load(open(gpath))
        if golden.get("fingerprint") != fingerprint(problem) or "converged_run" not in golden:
            log("golden mismatch: the generated problem differs from the golden one")
            golden = None
        else:
            gstate = dict(np.load(os.path.join(ROOT, "tests", "golden", "aerial_oracle_state.npz")))
    f_star = golden["converged_run"]["final_cost"] if golden else None

    def time_to_tolerance(points, tol=1e-6):
        if f_star is None:
            return None
        for it, cost, tt in points:
            if (cost - f_star) / f_star <= tol:
                return {"seconds": round(tt, 6), "iteration": it}
        return None

    _, tr_m = solve(PO.Precision.Mixed, trace=True)
    _, tr_64 = solve(PO.Precision.Fp64, trace=True)
    ttt_m, ttt_64 = time_to_tolerance(tr_m), time_to_tolerance(tr_64)
    log(f"time to 1e-6 of the converged oracle cost: mixed {ttt_m}, fp64 {ttt_64}")

    # the reference's GPU algorithm (FP64 block-Jacobi PCG to 2e-10) and the truncated-Newton FP32 PCG
    # configuration, over the reference's 20-iteration budget
    pcg = trunc = None
    if args.pcg_iterations > 0:
        rp, _ = solve(PO.Precision.Fp64, max_iterations=args.pcg_iterations, route=PO.ReducedSolver.Pcg)
        log(f"fp64 PCG ({args.pcg_iterations} it budget): {brief(rp)}")
        pcg = {"iterations": rp.report["iterations"], "solve_s": round(rp.report["solve_seconds"], 4),
               "lm_iter_per_s": round(rp.report["iterations"] / rp.report["solve_seconds"], 3),
               "pcg_iterations": rp.report["linear_iterations"], "final_cost": rp.finalCost}

        def truncated(o):
            o.backend.mixedSwitchDecrease = -1.0  # FP32 to the end

        rt, _ = solve(PO.Precision.Mixed, max_iterations=args.pcg_iterations, route=PO.ReducedSolver.Pcg,
                      configure=truncated)
        log(f"mixed truncated-Newton PCG ({args.pcg_iterations} it budget): {brief(rt)}")
        trunc = {"iterations": rt.report["iterations"], "solve_s": round(rt.report["solve_seconds"], 4),
                 "lm_iter_per_s": round(rt.report["iterations"] / rt.report["solve_seconds"], 3),
                 "pcg_iterations": rt.report["linear_iterations"], "final_cost": rt.finalCost,
                 "rel_cost_err_vs_converged_oracle": abs(rt.finalCost - f_star) / f_star if f_star else None}

    # adjoint at the converged mixed state: l = sum |C - T|^2, gradients for every observation
    adjoint = quality = None
    if not args.no_adjoint:
        r = mixed[-1]
        solve(PO.Precision.Mixed)
        T = r.cam_center + np.random.default_rng(5).normal(0.0, 0.05, r.cam_center.shape)
        t0 = time.perf_counter()
        g = ws.backward(problem, 2.0 * (r.cam_center - T))
        adjoint = {"seconds": round(time.perf_counter() - t0, 4), "loss": "sum |C - T|^2 over refined centres",
                   "state": "converged mixed solve", "gradient_entries": int(g.size),
                   "nonzero": int(np.count_nonzero(g)), "finite": bool(np.all(np.isfinite(g)))}
        log(f"adjoint: {adjoint}")
        qo = PO.quality_to_c(S.aerial_options(PO.SolveOptions()))
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            q = ws.quality(qo)
            times.append(time.perf_counter() - t0)
        quality = {"seconds": round(min(times), 5), "total_tracks": q["total_tracks"],
                   "optimized_tracks": q["optimized_tracks"], "valid_track_ratio": round(q["valid_track_ratio"], 5),
                   "mean_rms_before_px": round(q["mean_rms_before"], 5), "mean_rms_after_px": round(q["mean_rms_after"], 5),
                   "filter_threshold_px": round(q["filter_threshold"], 5),
                   "quality_gate_rejected": bool(q["quality_gate_rejected"])}
        log(f"quality: {quality}")

    mixed_t = sum(r.report["solve_seconds"] for r in mixed) / len(mixed)
    mixed_it = sum(r.report["iterations"] for r in mixed) / len(mixed)
    fp64_t = sum(r.report["solve_seconds"] for r in fp64s) / len(fp64s)
    fp64_it = sum(r.report["iterations"] for r in fp64s) / len(fp64s)

    # Roofline of the dominant kernel family of the timed mixed solve (average timed launch x launches). The direct
    # factorisations are MFMA work (dense: n^3 / 3 flops; band: the flops of the planned cyclic reduction; FP64
    # once the solve has moved to FP64 assembly), the other families move memory (algorithmic bytes).
    def family_line(st):
        out = {}
        for k, v in st.items():
            avg = v["total_ms"] / v["launches"] if v["launches"] else 0.0
            rate = v["bytes_per_launch"] / (avg * 1e-3) if avg > 0 else 0.0
            out[k] = {"launches": v["launches"], "avg_ms": round(avg, 6), "total_ms": round(v["total_ms"], 3),
                      ("TFLOPs" if k in MFMA_FAMILIES else "GBps"): round(rate / (1e12 if k in MFMA_FAMILIES else 1e9), 2)}
        return out

    fam = family_line(stats)
    dom = max(fam, key=lambda k: fam[k]["total_ms"])
    s_dom = stats[dom]
    avg_ms = fam[dom]["avg_ms"]
    if dom in MFMA_FAMILIES:
        achieved = s_dom["bytes_per_launch"] / (avg_ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 3), "peak": FP64_MFMA_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 5),
                    "flops_per_launch": s_dom["bytes_per_launch"]}
    else:
        achieved = s_dom["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "bytes_per_launch": s_dom["bytes_per_launch"]}
    roofline.update({"traffic": profiled_traffic(dom, fingerprint(problem)) if not args.small else None,
                     "avg_launch_ms": avg_ms, "timed_launches": s_dom["launches"], "families": fam,
                     "families_fp64_solve": family_line(stats64)})

    rel = lambda r: abs(r.finalCost - f_star) / f_star if f_star else None  # noqa: E731
    state_vs_oracle = None
    if gstate is not None:
        state_vs_oracle = {"oracle": "converged FP64 oracle (tests/golden/aerial_oracle_state.npz)",
                           "mixed": state_differences(gstate, mixed[0]), "fp64": state_differences(gstate, fp64)}
        log(f"state vs oracle: {state_vs_oracle}")

    batch = None
    if args.batch_windows > 0 and not args.small:
        try:
            precs = [PO.Precision.Mixed] + ([PO.Precision.Fp64] if args.batch_fp64 else [])
            batch = batch_leg(args.batch_windows, precs, args.batch_oracle_sample, bool(args.batch_warmup))
        except Exception as e:  # the batch leg reports its failure; the headline line still prints
            batch = {"error": f"{type(e).__name__}: {e}"}
            log(f"batch leg failed: {batch['error']}")

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        # the FP64 oracle on all host cores over the reference's 20-iteration budget (~7 s on 16 threads), and a
        # one-thread sample of its first iteration
        o = PO.to_c(S.aerial_options(PO.SolveOptions()))
        o.max_iterations = args.cpu_iterations
        threads = O.lib().oracle_max_threads()
        orc = O.Oracle(problem, threads=threads)
        t0 = time.perf_counter()
        rep = orc.solve(o)
        wall = time.perf_counter() - t0
        cpu = {"value": round(rep["iterations"] / rep["solve_seconds"], 4), "unit": "LM iter/s", "cores": threads,
               "kind": "port",
               "sample": f"FP64 oracle, first {rep['iterations']} LM iterations of the same workload "
                         f"({rep['solve_seconds']:.2f}s solve, {wall:.2f}s with setup); to converge it needs "
                         f"{golden['converged_run']['iterations'] if golden else '?'} iterations"}
        if args.cpu_single_iterations > 0:
            o1 = PO.to_c(S.aerial_options(PO.SolveOptions()))
            o1.max_iterations = args.cpu_single_iterations
            rep1 = O.Oracle(problem, threads=1).solve(o1)
            cpu["single_thread"] = {"value": round(rep1["iterations"] / rep1["solve_seconds"], 4), "cores": 1,
                                    "sample": f"first {rep1['iterations']} LM iterations "
                                              f"({rep1['solve_seconds']:.2f}s solve)"}
        log(f"cpu baseline: {cpu}")
        orc.close()

    out = {
        "metric": METRIC,
        "value": round(mixed_it / mixed_t, 4),
        "unit": "LM iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed_t * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (Jacobians/Schur until the FP32 rounding floor) + fp64 (cost, acceptance, reduced solve, the "
                 "converging steps)",
        "data": "synthetic",
        "config": {
            "workload": ("aerial block (reduced check size)" if args.small else
                         "aerial block: 1000 cameras in 3 strips at 300 m, 200k points, 1.5M observations, seed 2"),
            "observations": problem.num_observations,
            "active_observations": mixed[0].report["observation_count"],
            "cameras": problem.num_cameras,
            "points": problem.num_tracks,
            "loss": "Cauchy 1 px",
            "refined_intrinsics": "shared focal + k1,k2,k3,p1,p2 over 4 groups",
            "stop": f"step RMS <= 1e-6 (reference), at most {args.max_iterations} LM iterations",
            "path": f"online point Schur, reduced solve: {mixed[0].report['linear_solver_name']}",
        },
        "time_to_converge_s": round(mixed_t, 6),
        "lm_iterations": mixed_it,
        "mixed_fp32_accepted_steps_reported": bool(mixed[0].report["mixed_precision_used"]),
        "mixed_message": mixed[0].report["message"],
        "fp64_time_to_converge_s": round(fp64_t, 6),
        "fp64_lm_iterations": fp64_it,
        "fp64_lm_iter_per_s": round(fp64_it / fp64_t, 4),
        "speedup_mixed_vs_fp64": round(fp64_t / mixed_t, 4),
        "time_to_1e-6_cost": {"mixed": ttt_m, "fp64": ttt_64,
                              "speedup_mixed_vs_fp64": round(ttt_64["seconds"] / ttt_m["seconds"], 4)
                              if ttt_m and ttt_64 else None},
        "final_cost": mixed[0].finalCost,
        "rel_cost_err_vs_oracle": rel(mixed[0]),
        "fp64_rel_cost_err_vs_oracle": rel(fp64),
        "meets_1e-6": (rel(mixed[0]) <= 1e-6) if f_star else None,
        "oracle_converged": {"iterations": golden["converged_run"]["iterations"],
                             "final_cost": f_star} if golden else None,
        "fp32_seed_mode": {"time_to_converge_s": round(seed.report["solve_seconds"], 6),
                           "iterations": seed.report["iterations"], "seeds_accepted": seed.report["mixed_seed_steps"],
                           "seeds_rejected": seed.report["dense_fallbacks"],
                           "mixed_precision_used": seed.report["mixed_precision_used"],
                           "rel_cost_err_vs_oracle": rel(seed)},
        "fp64_pcg_reference_algorithm": pcg,
        "mixed_truncated_newton_pcg": trunc,
        "adjoint": adjoint,
        "quality": quality,
        "batch": batch,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "state_vs_oracle": state_vs_oracle,
        "device": mixed[0].report["device_name"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
