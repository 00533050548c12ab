"""Benchmark of the explicit RK path on one MI355X.

Workload: Navier-Stokes Taylor-Green vortex, Re 1600, Ma 0.1, [0,2pi]^3
periodic, 96^3 hexahedra Kuhn-split into 6 tetrahedra (5,308,416 cells),
FP64, AUSM + Green-Gauss + Barth-Jespersen, RK3-3.  A "step" is one RK3-3
time step (3 stages) over the whole mesh with the state resident in HBM.

Prints one JSON line: metric "cell-stage updates/s", the roofline of the
dominant kernel (algorithmic bytes per launch over its HIP-event-timed
average duration) and the CPU baseline (the oracle, one core, on a bounded
sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X published HBM3E peak

# Algorithmic bytes per cell and kernel launch (FP64, int32 indices, each
# array element once per launch), tetrahedra (2 faces per cell, 4 faces in a
# cell's list), second order, viscous.
#   gradient kernel: U read 40, face n/area/lambda/cells 48 x 2 faces,
#     cell->face list 4*4+4, 1/V 8, gradients write 240, min/max write 160
#   residual/stage kernel: U read 40 + write 40, RK register read + write 80,
#     face n/t1/t2/area/cells/x_f 112 x 2 faces, centroid 24, V 8, dx 24,
#     cell->face list 20, gradients read 240, min/max read 160
B_GRAD = 40 + 96 + 20 + 8 + 240 + 160
B_STAGE = 80 + 80 + 224 + 24 + 8 + 24 + 20 + 240 + 160
# Whole-stage compulsory traffic (each array element once per stage).
B_CELL_STAGE = 1276


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=96, help="hexahedra per side")
    ap.add_argument("--tile", type=int, default=8, help="hexahedra per side of a "
                    "cell-ordering tile of the mesh")
    ap.add_argument("--cpu-sample-n", type=int, default=64)
    ap.add_argument("--cpu-sample-steps", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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

    t0 = time.time()
    cfg = configs.taylor_green(n=args.n, steps=args.steps + args.warmup + 100, tile=args.tile)
    case = S.Case(**cfg)
    print("mesh + init: %.1f s, %d cells, %d faces" % (time.time() - t0, case.n_cells,
                                                    case.n_faces), file=sys.stderr, flush=True)
    sec, k_us = case.bench(args.steps, args.warmup)
    N = case.n_cells
    stages = 3
    value = N * stages * args.steps / sec
    ms_per_step = sec * 1e3 / args.steps

    k_grad_s, k_stage_s = k_us[0] * 1e-6, k_us[1] * 1e-6
    ach_stage = B_STAGE * N / k_stage_s / 1e9
    ach_grad = B_GRAD * N / k_grad_s / 1e9
    step_gbs = B_CELL_STAGE * N * stages * args.steps / sec / 1e9
    roofline = {"bound": "hbm", "kernel": "fv_residual_kernel<AUSM,linear,viscous>",
                "achieved": round(ach_stage, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach_stage / PEAK_HBM_GBS, 4), "traffic": None,
                "bytes_per_cell": B_STAGE, "avg_us": round(k_us[1], 2)}
    extra = {"gradient_kernel": {"achieved": round(ach_grad, 1), "bytes_per_cell": B_GRAD,
                                 "avg_us": round(k_us[0], 2),
                                 "frac": round(ach_grad / PEAK_HBM_GBS, 4)},
             "step": {"achieved": round(step_gbs, 1), "bytes_per_cell_stage": B_CELL_STAGE,
                      "frac": round(step_gbs / PEAK_HBM_GBS, 4)}}
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

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_sample_n, args.cpu_sample_steps)

    line = {
        "metric": "cell-stage updates/s (FP64 RK)",
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
        "config": {"workload": "taylor-green Re1600 Ma0.1, %d^3 hex x 6 tets (%d cells), "
                               "AUSM + Green-Gauss + Barth-Jespersen, RK3-3" % (args.n, N),
                   "cells": N, "stages_per_step": stages},
        "roofline": roofline,
        "roofline_detail": extra,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(n, steps):
    """The oracle (one core, the reference's data model) on a smaller mesh of
    the same workload; cell-stage updates per second."""
    import os as _os
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
    cpu_name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": N * 3 * done / sec, "unit": "cell-stage/s", "cores": 1, "kind": "port",
            "sample": "oracle (C99, reference AoS data model, gcc -O3) on the same workload "
                      "at %d^3 x 6 = %d cells, %d RK3-3 steps, %.1f s on %s (nproc %d)"
                      % (n, N, done, sec, cpu_name, _os.cpu_count() or 0)}


if __name__ == "__main__":
    main()
