"""Benchmark of the residual and time-stepping path on one MI355X.

Headline workload: Navier-Stokes Taylor-Green vortex, Re 1600, Ma 0.1,
[0,2pi]^3 periodic, 96^3 hexahedra Kuhn-split into 6 tetrahedra (5,308,416
cells), FP64, AUSM + Green-Gauss + Barth-Jespersen, RK3-3.  A "step" is one
RK3-3 time step (3 stages) over the whole mesh with the state resident in
HBM.  `value` is cell-stage updates per second of that step loop.

Beside it, in the same JSON line:
- `jfnk`: the steady bump channel (160x40x13 hexahedra x 6 tetrahedra =
  499,200 cells, Ma 0.5): wall time to the 1e-8 stop norm of plain FP64
  block-Jacobi JFNK and of the variants (FP32 Krylov; learned diagonal
  blocks; the multicolour block-ILU coupling; their combinations), each
  with its speedup over plain and the relative difference of its solution;
  `headline` names the learned-preconditioner mixed-precision variant
  (learned FP32 with the coupling) and its speedup; `converged_1e-12` the
  relative difference of the plain and headline solutions driven to stop
  norm 1e-12; `survey_outlet_limiter` a bounded plain run of the channel
  with the zero-gradient outlet and the limiter;
- `batched`: 64 periodic cases of log-uniform size (seeded), one device
  state, graph-replayed steps on the autotuned schedule (per-group
  launches, lane per face slot or face-centred, or persistent work-queue
  launches per class), the other schedules beside it: aggregate cell
  updates per second and the per-case step time inside the batched
  persistent launches, p50/p99;
- `roofline` of the explicit stage (the gradient and residual kernels that
  make one RK stage): 1,276 algorithmic bytes per cell-stage times the cells
  over the two kernels' HIP-event-timed average durations, with the measured
  HBM bytes of the two launches when a counter pass of the current kernel
  sources is recorded (profiles/explicit_tgv96_traffic.json); per-kernel
  figures with their share of the byte budget in `roofline_detail`; and
  `cpu_baseline` (the oracle, one core, on a bounded sample of the same
  workload).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X published HBM3E peak

# Algorithmic bytes per cell-stage (FP64, int32 indices, each array element
# once per stage, tetrahedra: 2 faces per cell, 4 faces in a cell's list),
# second order, viscous: U read + write 80, RK register read + write 80,
# face n, t1, t2, area, lambda, x_f, cells 2 x 120, V + dx 32, centroid 24,
# cell->face list 20, gradients write + read 480, min/max write + read 320.
B_CELL_STAGE = 1276
# The same terms split between the two kernels of a stage, each term counted
# in the first kernel that touches it:
#   gradient kernel: U read 40, face n/area/lambda/cells 2 x 48, cell->face
#     list 20, 1/V 8, gradients write 240, min/max write 160
#   residual/stage kernel: U write 40, RK register read + write 80, face
#     t1/t2/x_f 2 x 72, centroid 24, dx 24, gradients read 240, min/max
#     read 160
B_GRAD = 40 + 96 + 20 + 8 + 240 + 160
B_STAGE = 40 + 80 + 144 + 24 + 24 + 240 + 160
assert B_GRAD + B_STAGE == B_CELL_STAGE
# Tetrahedra, first order, inviscid (batched Euler cases)
B_CELL_STAGE_FO = 388
# Measured HBM bytes per launch of the two kernels (rocprofv3 FETCH_SIZE and
# WRITE_SIZE scaled by a device copy's counters), valid only for the kernel
# sources they were measured on (their SHA-256 is recorded with them).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "explicit_tgv96_traffic.json")
KERNEL_SOURCES = ["fv3d_amd/csrc/fv_kernels.hip", "fv3d_amd/csrc/fv_physics.h",
                  "fv3d_amd/csrc/fv_device.h"]


def kernel_source_hash():
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def measured_traffic(n):
    """Measured HBM bytes per launch {gradient, residual} for the current
    kernel sources on the n^3 x 6 mesh, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("sources_sha256") != kernel_source_hash() or t.get("n") != n:
        return None
    return t["bytes_per_launch"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=96, help="hexahedra per side")
    ap.add_argument("--tile", type=int, default=4, help="hexahedra per side of a "
                    "cell-ordering tile of the mesh")
    ap.add_argument("--cpu-sample-n", type=int, default=64)
    ap.add_argument("--cpu-sample-steps", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-explicit", action="store_true",
                    help="skip the headline explicit run (profiling the other legs)")
    ap.add_argument("--no-jfnk", action="store_true")
    ap.add_argument("--jfnk-mesh", default="160,40,13")
    ap.add_argument("--jfnk-max-newton", type=int, default=400,
                    help="Newton step limit of the JFNK legs (a short profiling run)")
    ap.add_argument("--jfnk-variants", default="",
                    help="comma-separated subset of the JFNK variants after the plain "
                         "FP64 block-Jacobi run (profiling)")
    ap.add_argument("--no-jfnk-agreement", action="store_true",
                    help="skip the JFNK runs to stop norm 1e-12 and the survey-outlet run")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-cases", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-alone", action="store_true",
                    help="also time every batched case alone")
    ap.add_argument("--batch-autotune", action="store_true",
                    help="re-run the batched autotuner instead of using its cache")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (torch) whose HBM "
                         "byte counters calibrate FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build runs the whole mesh on one GPU; runs on "
              "%d GPUs are not part of it" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from fv3d_amd import configs
    from fv3d_amd import solver as S

    value = ms_per_step = N = None
    roofline = extra = None
    stages = 3
    if not args.no_explicit:
        t0 = time.time()
        cfg = configs.taylor_green(n=args.n, steps=args.steps + args.warmup + 100, tile=args.tile)
        case = S.Case(**cfg)
        print("mesh + init: %.1f s, %d cells, %d faces" % (time.time() - t0, case.n_cells,
                                                        case.n_faces), file=sys.stderr, flush=True)
        sec, k_us = case.bench(args.steps, args.warmup)
        N = case.n_cells
        value = N * stages * args.steps / sec
        ms_per_step = sec * 1e3 / args.steps

        k_grad_s, k_stage_s = k_us[0] * 1e-6, k_us[1] * 1e-6
        # the residual kernel on tetrahedra: face-centred unless switched off
        k2 = ("fv_residual_face_kernel<AUSM,face states,viscous>" if os.environ.get("FVH_K2_FACE", "1") != "0"
              else "fv_residual_group_kernel<AUSM,face states,viscous,4>")
        traffic = measured_traffic(args.n)
        # roofline: the whole stage (gradient + residual kernel launches,
        # HIP-event timed on the library's stream) on the survey's bytes per
        # cell-stage
        stage_s = k_grad_s + k_stage_s
        ach = B_CELL_STAGE * N / stage_s / 1e9
        roofline = {
            "bound": "hbm",
            "kernel": "explicit stage = fv_gradient_tile_kernel<4, double> + " + k2,
            "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(ach / PEAK_HBM_GBS, 4),
            "traffic": (traffic["gradient"] + traffic["residual"]) if traffic else None,
            "bytes_per_cell_stage": B_CELL_STAGE, "cells": N,
            "stage_us": round(stage_s * 1e6, 2)}
        ach_stage = B_STAGE * N / k_stage_s / 1e9
        ach_grad = B_GRAD * N / k_grad_s / 1e9
        step_gbs = B_CELL_STAGE * N * stages * args.steps / sec / 1e9
        extra = {
            "residual_kernel": {
                "kernel": k2,
                "budget_bytes_per_cell": B_STAGE,
                "achieved": round(ach_stage, 1), "unit": "GB/s",
                "frac": round(ach_stage / PEAK_HBM_GBS, 4),
                "traffic": traffic["residual"] if traffic else None,
                "avg_us": round(k_us[1], 2)},
            "gradient_kernel": {
                "kernel": "fv_gradient_tile_kernel<4, double>",
                "budget_bytes_per_cell": B_GRAD,
                "achieved": round(ach_grad, 1), "unit": "GB/s",
                "frac": round(ach_grad / PEAK_HBM_GBS, 4),
                "traffic": traffic["gradient"] if traffic else None,
                "avg_us": round(k_us[0], 2)},
            "step": {"achieved": round(step_gbs, 1), "bytes_per_cell_stage": B_CELL_STAGE,
                     "frac": round(step_gbs / PEAK_HBM_GBS, 4),
                     "note": "timed step loop (all kernels, 3 stages per step)"}}
        print("explicit: %.3f ms/step, %.3e cell-stage/s" % (ms_per_step, value),
              file=sys.stderr, flush=True)
        case.close()

    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        print("device copy 1 GiB: %.3f ms, %.0f GB/s read+write" % (ms, 2 * (1 << 30) / ms / 1e6),
              file=sys.stderr, flush=True)
        del src, dst

    jfnk = None if args.no_jfnk else bench_jfnk([int(x) for x in args.jfnk_mesh.split(",")],
                                                 args.jfnk_max_newton,
                                                 args.jfnk_variants.split(",")
                                                 if args.jfnk_variants else None,
                                                 not args.no_jfnk_agreement)
    batched = None if args.no_batch else bench_batch(args.batch_cases, args.batch_steps,
                                                     args.batch_alone, args.batch_autotune)

    cpu = None
    if not args.no_cpu_baseline and not args.no_explicit:
        cpu = cpu_baseline(args.cpu_sample_n, args.cpu_sample_steps)
        if jfnk is not None:
            cpu["jfnk"] = cpu_baseline_jfnk()

    line = {
        "metric": "cell-stage updates/s (FP64 RK) + JFNK time-to-1e-8 residual, learned vs plain",
        "value": value,
        "unit": "cell-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "taylor-green Re1600 Ma0.1, %d^3 hex x 6 tets (%s cells), "
                               "AUSM + Green-Gauss + Barth-Jespersen, RK3-3"
                               % (args.n, N if N is not None else 6 * args.n ** 3),
                   "cells": N, "stages_per_step": stages},
        "roofline": roofline,
        "roofline_detail": extra,
        "jfnk": jfnk,
        "batched": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_jfnk(n, max_newton=400, only=None, agreement=True):
    """Plain FP64 block-Jacobi JFNK against the learned-preconditioner,
    FP32-Krylov JFNK on the bump channel, both from the free stream to the
    1e-8 stop norm."""
    import numpy as np
    from fv3d_amd import configs, learned
    from fv3d_amd import solver as S

    cfg = configs.bump_channel_steady(nx=n[0], ny=n[1], nz=n[2])
    case = S.Case(**cfg)
    U0 = case.state()
    out = {"workload": "bump channel %dx%dx%d hex x 6 tets (%d cells), Ma 0.5, AUSM, linear "
                       "reconstruction, slip walls, state outlet; stop norm 1e-8"
                       % (n[0], n[1], n[2], case.n_cells), "cells": case.n_cells}
    st, _ = case.jfnk(tol=1e-8, precision="fp64", precond="block", max_newton=max_newton)
    U64 = case.state()
    out["plain_fp64_block"] = _jfnk_row(st)
    print("jfnk plain: %.2f s, %d newton, %d gmres" % (st["seconds"], st["newton_steps"],
                                                        st["gmres_iterations"]),
          file=sys.stderr, flush=True)
    # GMRES restart length: the block-Jacobi runs are fastest at 30 (plain
    # FP64: 7.97 s, against 9.14 s at 40 and 10.3 s at 50); the coupled
    # preconditioner needs fewer iterations per Newton step and is fastest at
    # 40 (2.08 s against 2.19 s at 30 and 2.25 s at 50, learned FP32); both
    # allow at most 120 iterations per Newton step
    coupled = dict(coupling="dilu", krylov_dims=40, max_restarts=2)
    variants = [("fp32_block", dict(precision="fp32", precond="block")),
                ("fp64_dilu", dict(precision="fp64", precond="block", **coupled)),
                ("fp32_dilu", dict(precision="fp32", precond="block", **coupled))]
    have_net = os.path.exists(learned.PRETRAINED)
    if have_net:
        learned.load_pretrained()
        variants += [("learned_fp32", dict(precision="fp32", precond="learned")),
                     ("learned_fp32_dilu", dict(precision="fp32", precond="learned",
                                                **coupled))]
    else:
        out["learned_fp32"] = out["learned_fp32_dilu"] = None
        out["note"] = "no pretrained network at %s" % learned.PRETRAINED
    if only is not None:
        variants = [v for v in variants if v[0] in only]
    for name, kw in variants:
        case.set_state(U0)
        st, _ = case.jfnk(tol=1e-8, max_newton=max_newton, **kw)
        U = case.state()
        row = _jfnk_row(st)
        row["solution_rel_diff_vs_plain"] = float(np.linalg.norm(U - U64) / np.linalg.norm(U64))
        row["speedup_vs_plain"] = out["plain_fp64_block"]["seconds"] / st["seconds"]
        out[name] = row
        print("jfnk %s: %.2f s, %d newton, %d gmres, fallback %d" % (
            name, st["seconds"], st["newton_steps"], st["gmres_iterations"],
            st["precond_fallbacks"]), file=sys.stderr, flush=True)
    if agreement and have_net and (only is None or "learned_fp32_dilu" in only):
        # the converged solutions (stop norm 1e-12) of plain FP64 block
        # Jacobi and of the learned FP32 + DILU variant
        rows = {}
        sol = {}
        for name, kw in (("plain_fp64_block", dict(precision="fp64", precond="block")),
                         ("learned_fp32_dilu", dict(precision="fp32", precond="learned",
                                                    **coupled))):
            case.set_state(U0)
            st, _ = case.jfnk(tol=1e-12, max_newton=max_newton, **kw)
            rows[name] = _jfnk_row(st)
            sol[name] = case.state()
        rows["solution_rel_diff"] = float(np.linalg.norm(sol["learned_fp32_dilu"] -
                                                         sol["plain_fp64_block"]) /
                                          np.linalg.norm(sol["plain_fp64_block"]))
        out["converged_1e-12"] = rows
        print("jfnk converged to 1e-12: rel diff %.3g" % rows["solution_rel_diff"],
              file=sys.stderr, flush=True)
    case.close()
    if agreement:
        # the channel as the survey states it: zero-gradient outlet and the
        # Barth-Jespersen limiter (plain FP64 block Jacobi, bounded run)
        cfg_s = configs.bump_channel(nx=n[0], ny=n[1], nz=n[2])
        case = S.Case(**cfg_s)
        st, _ = case.jfnk(tol=1e-8, precision="fp64", precond="block", max_newton=150)
        out["survey_outlet_limiter"] = dict(_jfnk_row(st), note=(
            "zero-gradient outlet + Barth-Jespersen limiter; status 1 = the Newton step "
            "limit (150) was reached before the 1e-8 stop norm"))
        print("jfnk survey outlet/limiter: status %d, norm %.3g after %d newton" % (
            st["status"], st["norm"], st["newton_steps"]), file=sys.stderr, flush=True)
        case.close()
    best = out.get("learned_fp32_dilu")
    if best:
        out["headline"] = {"variant": "learned_fp32_dilu",
                           "speedup_vs_plain": best["speedup_vs_plain"],
                           "seconds": best["seconds"],
                           "plain_seconds": out["plain_fp64_block"]["seconds"]}
    return out


def _jfnk_row(st):
    return {"seconds": st["seconds"], "status": st["status"], "newton": st["newton_steps"],
            "gmres": st["gmres_iterations"], "residual_evals": st["residual_evals"],
            "norm": st["norm"], "l2_drop": st["l2_drop"],
            "precond_seconds": st["precond_seconds"],
            "precond_fallbacks": st["precond_fallbacks"]}


def bench_batch(n_cases, steps, alone=False, retune=False):
    """The batched cases on the schedule the autotuner picked (cached in
    fv3d_amd/data/batch_tuning.json for this case set; tuned here when the
    cache is missing or --batch-autotune): aggregate cell updates per second
    of graph-replayed steps, the per-group schedule beside it, and the
    per-case step time inside the batched launches (p50/p99)."""
    import numpy as np
    from fv3d_amd import batch as B

    cases = B.c5_cases(n_cases)
    tuned = None if retune else B.load_tuning(cases)
    tuned_here = False
    if tuned is None:
        tuned = B.autotune(cases)
        B.save_tuning(cases, tuned[0], tuned[1], device="MI355X")
        tuned_here = True
    best, table = tuned
    sched = dict(persistent=best["persistent"], block=best["block"] or 256,
                 grid=best["grid"], sub_tiles=best.get("sub_tiles") or 1,
                 face=bool(best.get("face", False)))
    pers = [r for r in table if r["persistent"]]
    best_p = min(pers, key=lambda r: r["ms_per_step"]) if pers else None
    t0 = time.time()
    b = B.Batch(cases, tile=best["tile"], schedule=sched)
    setup = time.time() - t0
    b.run(3, graph=False)
    b.run(2, graph=True)
    sec = b.run(steps, graph=True)
    sec_eager = b.run(steps, graph=False)
    b.set_schedule(False)
    b.run(1, graph=True)
    sec_group = b.run(steps, graph=True)
    b.set_schedule(False, face=True)
    b.run(1, graph=True)
    sec_face = b.run(steps, graph=True)
    # per-case step times inside the shared launches, from the tiles'
    # timestamps: on the face schedule, and on the persistent schedule's best
    # autotuned configuration
    per_face = b.case_step_times(5)
    per, sec_pers, p_sched = None, None, None
    if best_p is not None:
        p_sched = dict(persistent=True, block=best_p["block"], grid=best_p["grid"],
                       sub_tiles=best_p.get("sub_tiles") or 1)
        b.set_schedule(**p_sched)
        b.run(1, graph=True)
        sec_pers = b.run(steps, graph=True)
        per = b.case_step_times(5)
    cells = b.n_cells
    ms = sec / steps * 1e3
    out = {"workload": "%d periodic cases, 2e4-4e5 tets log-uniform (seeds 0..%d), Mach 0.2-2.0, "
                       "Rusanov/AUSM by seed parity, Navier-Stokes every 4th, first order, RK3-3"
                       % (n_cases, n_cases - 1),
           "cells": cells, "cases": n_cases, "setup_s": round(setup, 1),
           "schedule": dict(sched, mesh_tile=best["tile"],
                            source="autotuned here" if tuned_here else
                            "autotuner cache fv3d_amd/data/batch_tuning.json"),
           "ms_per_step_graph": ms, "ms_per_step_eager": sec_eager / steps * 1e3,
           "ms_per_step_group_schedule_graph": sec_group / steps * 1e3,
           "ms_per_step_face_schedule_graph": sec_face / steps * 1e3,
           "ms_per_step_persistent_schedule_graph": (sec_pers / steps * 1e3
                                                     if sec_pers else None),
           "persistent_schedule": p_sched,
           "cell_updates_per_s": cells * steps / sec,
           "cell_stage_updates_per_s": 3 * cells * steps / sec,
           "roofline_first_order_frac": round(B_CELL_STAGE_FO * 3 * cells * steps / sec / 1e9
                                              / PEAK_HBM_GBS, 4)}
    def pct(x, label):
        return {"p50": float(np.percentile(x, 50)), "p99": float(np.percentile(x, 99)),
                "max": float(x.max()), "schedule": label}
    # the headline schedule's per-case times, and the other timed schedule's
    if sched["persistent"] and per is not None:
        out["per_case_step_ms_in_batch"] = pct(per, "persistent, as persistent_schedule")
        out["per_case_step_ms_in_batch_face_schedule"] = pct(per_face, "face")
    else:
        out["per_case_step_ms_in_batch"] = pct(per_face, "face")
        if per is not None:
            out["per_case_step_ms_in_batch_persistent_schedule"] = pct(
                per, "persistent, as persistent_schedule")
    b.close()
    print("batched: %.3f ms/step (%d cells), group schedule %.3f ms, face schedule %.3f ms, "
          "persistent %s ms, in-batch per-case p50 %.4f p99 %.4f ms (%s schedule)"
          % (ms, cells, sec_group / steps * 1e3, sec_face / steps * 1e3,
             out["ms_per_step_persistent_schedule_graph"],
             out["per_case_step_ms_in_batch"]["p50"], out["per_case_step_ms_in_batch"]["p99"],
             out["per_case_step_ms_in_batch"]["schedule"].split(",")[0]),
          file=sys.stderr, flush=True)
    if not alone:
        return out
    # per-case step time, each case alone (graph replay)
    per = []
    for c in cases:
        bb = B.Batch([c], tile=best["tile"], schedule=sched)
        bb.run(2, graph=False)
        bb.run(1, graph=True)
        per.append(bb.run(10, graph=True) / 10 * 1e3)
        bb.close()
    per = np.array(per)
    out["per_case_step_ms_alone"] = {"p50": float(np.percentile(per, 50)),
                                     "p99": float(np.percentile(per, 99)),
                                     "sum": float(per.sum())}
    return out


def _cpu_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, steps):
    """The oracle (one core, the reference's data model) on a smaller mesh of
    the same workload; cell-stage updates per second."""
    from fv3d_amd import configs
    from fv3d_amd import solver as S
    from oracle import fvref

    cfg = configs.taylor_green(n=n, steps=steps)
    case = S.Case(**cfg)
    mesh = S.mesh_arrays()
    U0 = case.state()
    case.close()
    o = fvref.from_case(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    o.set_state(U0.reshape(-1))
    t0 = time.perf_counter()
    done, _ = o.timedisc("RK3-3", steps, True, 1e-10, 0.0, 1e300)
    sec = time.perf_counter() - t0
    N = mesh["n_cells"]
    return {"value": N * 3 * done / sec, "unit": "cell-stage/s", "cores": 1, "kind": "port",
            "sample": "oracle (C99, reference AoS data model, gcc -O3) on the same workload "
                      "at %d^3 x 6 = %d cells, %d RK3-3 steps, %.1f s on %s (nproc %d)"
                      % (n, N, done, sec, _cpu_name(), os.cpu_count() or 0)}


def cpu_baseline_jfnk(n=(16, 4, 3)):
    """The oracle's JFNK (numpy driver over the C residual, one core) on a
    small mesh of the same channel: seconds to the 1e-8 stop norm."""
    import numpy as np
    from fv3d_amd import configs
    from fv3d_amd import solver as S
    from fv3d_amd._lib import lib
    from oracle.jfnk import OracleJfnk

    cfg = configs.bump_channel_steady(nx=n[0], ny=n[1], nz=n[2])
    L = lib()
    L.init_solver(b"cpu-baseline")
    S.check(L.fvh_mesh_channel(n[0], n[1], n[2], 0.05, 20261203))
    mesh = S.mesh_arrays()
    oj = OracleJfnk(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    u0 = 0.5 * math.sqrt(1.4)
    U = np.zeros((mesh["n_cells"], 5))
    U[:, 0] = 1.0
    U[:, 1] = u0
    U[:, 4] = 1 / 0.4 + 0.5 * u0 * u0
    _, st, _ = oj.solve(U, tol=1e-8)
    return {"seconds": st["seconds"], "cells": mesh["n_cells"], "newton": st["newton_steps"],
            "gmres": st["gmres_iterations"], "status": st["status"], "cores": 1, "kind": "port",
            "sample": "oracle JFNK (numpy GMRES over the C99 residual, FP64 block Jacobi) on the "
                      "%dx%dx%d channel, %s" % (n[0], n[1], n[2], _cpu_name())}


if __name__ == "__main__":
    main()
