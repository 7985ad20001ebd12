#!/usr/bin/env python3
"""Time steps/s of the MHFEM per-time-step path on one MI355X.

Workload: 3-D Darcy-type pressure on the unit cube, 64^3 cubes x 6 Kuhn
tetrahedra (1,572,864 cells, 3,170,304 faces = DOFs), n=1, N=0.2, m=1,
D = K with ln K a seeded Gaussian field (exponential covariance, sigma=1,
correlation length 0.1), dt=1e-3, BiCGStab to 1e-10 relative residual.
A "step" is one time step: preIterate + assembleLinearSystem +
solveLinearSystem + postIterate (control.h:74-101), all inputs resident in HBM.

Two forms are timed on the same workload from the same initial state:
  plain  -- FP64 BiCGStab + ILU(0), the configured reference method;
  chosen -- the fastest form that still stops at 1e-10 relative residual in
            FP64 (FP32 inner BiCGStab inside FP64 iterative refinement).
`value` is the chosen form's steps/s; the plain form is reported beside it.

roofline: the SpMV (the kernel that dominates the chosen form) is timed live
with HIP events on the library's stream; achieved = algorithmic bytes per
launch / average launch time.  cpu_baseline: the C oracle (a restatement of
the reference host path) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tnl-mhfem_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "time steps/s at 1e-10 rel. residual, chosen vs plain GPU form; HBM GB/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(rows, nnz, fp32):
    # algorithmic bytes of one CSR SpMV with int32 columns: values + columns,
    # row pointers, x read once, y written once
    if fp32:
        return 8 * nnz + 4 * (rows + 1) + 8 * rows
    return 12 * nnz + 4 * (rows + 1) + 16 * rows


def time_form(pr, form, steps, warmup):
    from mhfem_amd import capi

    s = capi.Solver(pr, **form["solver"])
    t = 0.0
    for _ in range(warmup):
        s.step(t, pr.dt)
        t += pr.dt
    s.synchronize()
    its0 = len(s.iterations)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(t, pr.dt)
        t += pr.dt
    s.synchronize()
    dt = time.perf_counter() - t0
    its = s.iterations[its0:]
    return s, dt, its


def cpu_baseline(pr, sample_steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    threads = min(16, len(os.sched_getaffinity(0)))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = pyoracle.Oracle(pr)
    t = 0.0
    t0 = time.perf_counter()
    its = []
    for _ in range(sample_steps):
        it, _ = o.step(t, pr.dt)
        its.append(it)
        t += pr.dt
    dt = time.perf_counter() - t0
    return {
        "value": round(sample_steps / dt, 5),
        "unit": "time steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {sample_steps} time step(s) of the same 64^3 Darcy workload, FP64 BiCGStab+ILU(0) "
                  f"to 1e-10 ({its} iterations), C oracle with OpenMP on {threads} host threads "
                  f"(ILU(0) triangular solves sequential)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=64, help="cubes per edge")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--chosen-prec", default="diagonal")
    ap.add_argument("--inner-tol", type=float, default=1e-3)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it "
                                   "(the sharding design is described in DESIGN.md)", "n_gpus": args.gpus}))
        return 2

    from mhfem_amd import problems

    log(f"[bench] building the {args.size}^3 Darcy workload")
    pr = problems.darcy_3d(s=args.size, seed=args.seed, preconditioner="ilu0")
    rows = pr.num_faces * pr.n

    chosen = {"name": f"FP32 inner BiCGStab+{args.chosen_prec} in FP64 refinement",
              "solver": dict(preconditioner=args.chosen_prec, precision="fp32-refined",
                             inner_tol=args.inner_tol, check_every=8, tol=1e-10)}
    plain = {"name": "FP64 BiCGStab+ILU(0)",
             "solver": dict(preconditioner="ilu0", precision="fp64", check_every=4, tol=1e-10)}

    log(f"[bench] chosen form: {chosen['name']}")
    s, dt, its = time_form(pr, chosen, args.steps, args.warmup)
    ms_step = 1e3 * dt / args.steps
    rp, col, val, rhs = s.export_system()
    nnz = int(rp[-1])
    spmv_ms = s.benchmark_spmv("fp32-refined", reps=50)
    spmv64_ms = s.benchmark_spmv("fp64", reps=50)
    log(f"[bench] chosen: {ms_step:.2f} ms/step, iterations {its}; spmv fp32 {spmv_ms*1e3:.1f} us, "
        f"fp64 {spmv64_ms*1e3:.1f} us")
    s.close()

    result_plain = None
    if not args.skip_plain:
        log(f"[bench] plain form: {plain['name']}")
        sp, dtp, itsp = time_form(pr, plain, args.steps, args.warmup)
        result_plain = {
            "form": plain["name"],
            "steps_per_s": round(args.steps / dtp, 4),
            "ms_per_step": round(1e3 * dtp / args.steps, 3),
            "iterations_per_step": round(sum(itsp) / len(itsp), 2),
        }
        log(f"[bench] plain: {result_plain}")
        sp.close()

    achieved = spmv_bytes(rows, nnz, True) / (spmv_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(args.steps / dt, 4),
        "unit": "time steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement",
        "data": "synthetic (seeded Gaussian ln K field, structured Kuhn mesh)",
        "config": {
            "workload": f"3-D Darcy pressure, {args.size}^3 cubes x 6 Kuhn tets, n=1, dt=1e-3, 1e-10 rel. residual",
            "cells": pr.num_cells, "faces": pr.num_faces, "dofs": rows, "nnz": nnz,
            "form": chosen["name"],
        },
        "iterations_per_step": round(sum(its) / len(its), 2),
        "dof_per_s": round(rows * args.steps / dt, 1),
        "plain": result_plain,
        "chosen_vs_plain": (round((args.steps / dt) / result_plain["steps_per_s"], 3) if result_plain else None),
        "roofline": {
            "kernel": "k_spmv<float> (SELL-64 SpMV of the FP32 inner solve)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "avg_launch_us": round(spmv_ms * 1e3, 2),
            "fp64_spmv_GBs": round(spmv_bytes(rows, nnz, False) / (spmv64_ms * 1e-3) / 1e9, 1),
        },
        "cpu_baseline": None,
    }
    if not args.skip_cpu:
        log("[bench] cpu baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(problems.darcy_3d(s=args.size, seed=args.seed, preconditioner="ilu0"),
                                           args.cpu_steps)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
