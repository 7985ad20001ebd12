#!/usr/bin/env python3
"""Time steps/s of the MHFEM per-time-step path on one MI355X.

Workload: n=2 coupled advection-diffusion-reaction on the unit cube, 64^3
cubes x 6 Kuhn tetrahedra (1,572,864 cells, 3,170,304 faces, 6,340,608
DOFs), D11=D22=K, D12=D21=0.1K with ln K a seeded Gaussian field
(exponential covariance, sigma=1, correlation length 0.1), a_ii=(1,.5,.25)
explicit upwind, r=[[1e-2,-5e-3],[-5e-3,1e-2]], Dirichlet (1, 0.5) at x=0,
dt=1e-3, every solve to 1e-10 relative residual.  A "step" is one time step:
preIterate + assembleLinearSystem + solveLinearSystem + postIterate
(control.h:74-101), all inputs resident in HBM.  The trajectory is the
workload's 50 steps: `warmup` untimed steps, then `steps` timed ones.

Forms, each run from the same initial state on a fresh solver:
  chosen     -- starting vector = previous solution + the least-residual
                combination of the last 16 increments, FP32 BiCGStab whose
                residual is replaced by the FP64 one every 33x drop (FP64
                iterative refinement with reliable updates), Jacobi (the
                identity on the unit-diagonal rows).  `value` is its steps/s.
  plain      -- FP64 BiCGStab + Jacobi, previous solution as initial guess:
                the fastest plain FP64 form on the GPU.
  plain_ilu0 -- FP64 BiCGStab + ILU(0), the reference's configured method
                (bounded sample: its triangular solves make a step take
                seconds).
  config2_as_configured -- FP32 BiCGStab + ILU(0) in FP32 inside FP64
                refinement (reliable updates), as configuration 2 names it.
  config1    -- n = 1 Darcy on the same mesh and K field, FP64 BiCGStab +
                ILU(0) (configuration 1), with the Jacobi form beside it.
  with_network -- the chosen form plus the learned network's direction
                (per-face MLP trained online on the first steps, switched off
                by its rejection rule): what the trained network adds.
  fp64_history -- FP64 BiCGStab + Jacobi from the previous increments: what
                the FP32 iteration adds, and (its distance to `plain`) the
                spread between two FP64 solutions that both meet 1e-10.

ensemble: the batched form -- 96 independent 2-D heat-type instances with
side s drawn from [32, 256] squares (log-normal K, sigma=1.5, correlation
length 0.05), FP64 BiCGStab + Jacobi to 1e-10, solved as one segmented system
with per-instance convergence; eager steps against steps replayed from
captured HIP graphs, after the rows-per-work-group autotuner, and against the
same instances stepped one after another with one context each.

roofline: the dominant kernel is the FP32 block-SELL SpMV of the chosen
form's iteration.  Every 8th launch during the chosen form's timed steps is
bracketed by HIP events on the library's stream (timing level 2);
achieved = the bytes the block-SELL product must move per launch (blocks of
n*n values with one column index, the faces' own off-diagonal entries, x
read once, y written once, and the operand of the dot product the kernel
fuses, read once) / average launch duration.  The same product
priced as an int32 CSR SpMV is reported beside it (csr_equivalent), and the
plain FP64 form's SpMV the same way (plain_spmv_fp64).
assembly: the same for assembleLinearSystem (one launch: interior and boundary rows),
timed with HIP events around every assembly of the chosen form's timed steps.
cpu_baseline: the C oracle (a restatement of the reference host path), the
first time step of the same workload, whole: FP64 BiCGStab + ILU(0) to 1e-10.
Its solution is the all-FP64 reference of error_vs_oracle_first_step.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tnl-mhfem_amd"))

HBM_PEAK_GBS = 8000.0
# HBM bytes per SpMV launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (scaled by a plain copy's reading); the summary holds the commands
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "spmv_block_sell_n2_cube64_hbm_counters.md")
METRIC = "time steps/s at 1e-10 rel. residual, chosen vs plain GPU form; HBM GB/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(rows, nnz, fp32):
    """Algorithmic bytes of one SpMV: CSR with int32 columns -- values and
    columns, row pointers, x read once, y written once."""
    if fp32:
        return 8 * nnz + 4 * (rows + 1) + 8 * rows
    return 12 * nnz + 4 * (rows + 1) + 16 * rows


def block_sell_bytes(n, faces, blocks_per_face, fp32, dot_operands=0.0):
    """Bytes the block-SELL SpMV must move for the same product: per stored
    off-diagonal block n*n values and one column index, per face the n(n-1)
    off-diagonal entries of its own block, x read once, y written once, and
    `dot_operands` vectors the fused dot products read beyond x.  The
    iteration's two launches fuse their dots into the product; the second
    product's operand s is its own input x, read once: the FP64 iteration's
    read r* and nothing more (0.5 on average over the alternating samples),
    the merged FP32 iteration's r* in both (1)."""
    w = 4 if fp32 else 8
    rows = n * faces
    return int(blocks_per_face * faces * (n * n * w + 4) + faces * n * (n - 1) * w + 2 * rows * w
               + dot_operands * rows * w)


def assembly_bytes(n, fpc, cells, faces, nnz):
    """Algorithmic bytes of one assembly (both the interior and the boundary
    rows): matrix values and right-hand side written, every cell's inputs
    read once (b_ijK_storage with its row sums, R_ijKe, R_iK, w and the
    right-hand-side terms), the face-cell and cell-face incidence."""
    S = 9 if fpc == 3 else 14
    rows = n * faces
    return 8 * nnz + 8 * rows + cells * 8 * (n * n * (S + fpc) + n * (2 + fpc)) + 4 * faces * 2 + 4 * cells * fpc


def profiled_traffic():
    """bytes per launch from the line 'traffic_bytes_per_launch: N' of the
    counter summary, None when it is absent"""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1].strip().split()[0])
    except OSError:
        pass
    return None


def run_form(pr, form, steps, warmup, timing=0):
    from mhfem_amd import capi

    s = capi.Solver(pr, **form["solver"])
    if form.get("guess"):
        s.configure_guess(True, **form["guess"])
    t = 0.0
    for w in range(warmup):
        s.step(t, pr.dt)
        t += pr.dt
        if w == 0:
            # the solution after the first step, for the comparison with the oracle's first step
            form["z_first_step"] = s.get_array("Z_iF").copy()
    s.synchronize()
    if form.get("autotune"):
        form["rows_per_block"] = s.autotune()[0]
    if timing:
        s.set_timing(timing)
        s.reset_timers()
    its0 = len(s.iterations)
    form["fallback_steps"] = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        form["fallback_steps"] += s.step(t, pr.dt).fallback_iterations > 0
        t += pr.dt
    s.synchronize()
    dt = time.perf_counter() - t0
    return s, dt, s.iterations[its0:]


def summary(name, steps, dt, its, rpb=None, fallback_steps=0):
    return {
        "form": name,
        "rows_per_block": rpb,
        "steps_with_fp64_fallback": fallback_steps,
        "steps_per_s": round(steps / dt, 4),
        "ms_per_step": round(1e3 * dt / steps, 3),
        "iterations_per_step": round(sum(its) / max(1, len(its)), 2),
    }


def ensemble_leg(steps, warmup, seed=7):
    from mhfem_amd import capi, problems

    sizes = problems.ensemble_sizes(seed, 96)
    members = [problems.ensemble_member(int(s_), seed * 1000 + k) for k, s_ in enumerate(sizes)]
    pr = problems.concatenate(members)
    out = {"instances": 96, "cells": pr.num_cells, "dofs": pr.num_faces * pr.n,
           "form": "FP64 BiCGStab + Jacobi, one segmented system"}
    # the alternative: one context per instance, stepped one after another
    solvers = [capi.Solver(m, preconditioner="diagonal", tol=1e-10, check_every=8) for m in members]
    t = 0.0
    for s in solvers:
        s.step(t, pr.dt)
    t += pr.dt
    for s in solvers:
        s.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in solvers:
            s.step(t, pr.dt)
        t += pr.dt
    for s in solvers:
        s.synchronize()
    dt = time.perf_counter() - t0
    for s in solvers:
        s.close()
    out["one_context_per_instance"] = {"steps_per_s": round(steps / dt, 3), "ms_per_step": round(1e3 * dt / steps, 3)}
    for mode in ("eager", "graph"):
        s = capi.Solver(pr, preconditioner="diagonal", tol=1e-10, check_every=8)
        rpb, _ = s.autotune()
        step = s.step if mode == "eager" else s.step_graph
        t = 0.0
        for _ in range(warmup):
            step(t, pr.dt)
            t += pr.dt
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(t, pr.dt)
            t += pr.dt
        s.synchronize()
        dt = time.perf_counter() - t0
        its, res = s.batch_stats()
        out[mode] = {"steps_per_s": round(steps / dt, 3), "ms_per_step": round(1e3 * dt / steps, 3),
                     "rows_per_block": rpb, "max_instance_iterations_last_step": int(its.max()),
                     "min_instance_iterations_last_step": int(its.min()),
                     "max_residue_last_step": float(res.max())}
        s.close()
    out["graph_vs_eager"] = round(out["graph"]["steps_per_s"] / out["eager"]["steps_per_s"], 3)
    out["graph_vs_one_context_per_instance"] = round(
        out["graph"]["steps_per_s"] / out["one_context_per_instance"]["steps_per_s"], 3)
    return out


def study96(size, train_seeds, train_steps, eval_seed, eval_steps, history_depth):
    """The learned-guess study on the large configuration: n = 2 on size^3
    cubes x 6 tets with the time-dependent Dirichlet value
    Z_1 = sin(2 pi t / 0.5) at x = 0.  The network is trained online on the
    first `train_steps` steps of each training K field, then evaluated with
    frozen weights on a held-out K field against the plain FP64 form and the
    method without the network."""
    from mhfem_amd import capi, problems

    mixed = dict(preconditioner="diagonal", precision="fp32-reliable", inner_tol=0.1, check_every=8, tol=1e-10)
    out = {"workload": f"n=2, {size}^3 cubes x 6 tets, Z_1 = sin(2 pi t/0.5) at x=0, dt=1e-3, 1e-10",
           "train_seeds": train_seeds, "train_steps": train_steps, "eval_seed": eval_seed,
           "eval_steps": eval_steps}
    w = None
    for sd in train_seeds:
        pr = problems.adr_3d(s=size, seed=sd, preconditioner="diagonal", time_dependent=True)
        s = capi.Solver(pr, **mixed)
        s.configure_guess(True, train_steps=train_steps, epochs=10, learning_rate=2e-3, seed=3, sample_stride=8,
                          history_depth=history_depth)
        if w is not None:
            s.set_guess_weights(w)
        t = 0.0
        for _ in range(train_steps):
            s.step(t, pr.dt)
            t += pr.dt
        w = s.guess_weights()
        s.close()
        log(f"[study] trained on seed {sd}")
    pr = problems.adr_3d(s=size, seed=eval_seed, preconditioner="diagonal", time_dependent=True)
    out["dofs"] = pr.num_faces * pr.n
    forms = {
        "plain_fp64": (dict(preconditioner="diagonal", check_every=8, tol=1e-10), None),
        "frozen_network_and_history": (mixed, dict(train_steps=0, combine="combined", history_depth=history_depth)),
        "history_only": (mixed, dict(train_steps=0, combine="history", history_depth=history_depth)),
        "fp64_history": (dict(preconditioner="diagonal", check_every=8, tol=1e-10),
                         dict(train_steps=0, combine="history", history_depth=history_depth)),
    }
    for name, (kw, guess) in forms.items():
        s = capi.Solver(pr, **kw)
        if guess:
            s.configure_guess(True, **guess)
            if guess["combine"] == "combined":
                s.set_guess_weights(w)
        t = 0.0
        s.step(t, pr.dt)
        t += pr.dt
        s.synchronize()
        s.autotune()
        t0 = time.perf_counter()
        its = []
        coefs = []
        fallbacks = 0
        for _ in range(eval_steps):
            st = s.step(t, pr.dt)
            its.append(st.iterations)
            fallbacks += st.fallback_iterations > 0
            if guess:
                coefs.append(s.guess_stats()["coef_network"])
            t += pr.dt
        s.synchronize()
        dt = time.perf_counter() - t0
        out[name] = {"steps_per_s": round(eval_steps / dt, 4), "ms_per_step": round(1e3 * dt / eval_steps, 2),
                     "iterations_per_step": round(sum(its) / len(its), 1), "steps_with_fp64_fallback": fallbacks,
                     "iterations": its}
        if guess and guess["combine"] == "combined":
            out[name]["mean_abs_network_coefficient"] = round(sum(abs(c) for c in coefs) / len(coefs), 5)
        log(f"[study] {name}: {out[name]}")
        s.close()
    for name in forms:
        if name != "plain_fp64":
            out[name]["vs_plain"] = round(out[name]["steps_per_s"] / out["plain_fp64"]["steps_per_s"], 3)
    return out


def cpu_baseline(pr, sample_iters):
    """The first time step of the workload on the C oracle, whole and timed:
    preIterate, assembly, FP64 BiCGStab + ILU(0) to 1e-10 (the reference's
    configured method; the ILU(0) factorization and triangular solves are
    sequential), postIterate, on the host's cores.  Returns the record and
    the oracle's Z_iF after the step (the all-FP64 reference the GPU forms'
    first steps are compared with).  A one-thread figure is extrapolated from
    `sample_iters` iterations of the same step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    # the GPU box's host share for one GPU is 16 cores
    threads = min(16, len(os.sched_getaffinity(0)))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    L = pyoracle.lib()
    L.or_set_threads(threads)
    o = pyoracle.Oracle(pr)
    it, res = C.c_int(0), C.c_double(0)
    t0 = time.perf_counter()
    o.pre_iterate(0.0, pr.dt)
    o.assemble(0.0, pr.dt)
    t_pre_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    conv = L.or_solve(o.p, pyoracle.SOLVERS["bicgstab"], pyoracle.PRECONDITIONERS["ilu0"], 1e-10, 100000, 30,
                      C.byref(it), C.byref(res))
    t_solve = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.post_iterate(0.0, pr.dt)
    t_post = time.perf_counter() - t0
    step_s = t_pre_asm + t_solve + t_post
    z = o.array("Z_iF").copy()
    iters = it.value
    del o

    def solve(orc, maxit):
        t0 = time.perf_counter()
        L.or_solve(orc.p, pyoracle.SOLVERS["bicgstab"], pyoracle.PRECONDITIONERS["ilu0"], 1e-10, maxit, 30,
                   C.byref(it), C.byref(res))
        return time.perf_counter() - t0

    # the same first step on one thread: the iteration time from a sample
    L.or_set_threads(1)
    o1 = pyoracle.Oracle(pr)
    t0 = time.perf_counter()
    o1.pre_iterate(0.0, pr.dt)
    o1.assemble(0.0, pr.dt)
    t1_pre_asm = time.perf_counter() - t0
    s1 = solve(o1, 1)
    s2 = solve(o1, 1 + sample_iters)
    t1_it = (s2 - s1) / sample_iters
    t1_factor = max(0.0, s1 - t1_it)
    t0 = time.perf_counter()
    o1.post_iterate(0.0, pr.dt)
    t1_post = time.perf_counter() - t0
    step1_s = t1_pre_asm + t1_factor + iters * t1_it + t1_post
    L.or_set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    rec = {
        "value": round(1.0 / step_s, 6),
        "unit": "time steps/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": model,
        "converged": int(conv),
        "iterations": iters,
        "one_thread": {"value": round(1.0 / step1_s, 6), "unit": "time steps/s",
                       "sample": (f"the same first step on 1 thread, extrapolated: preIterate+assembly "
                                  f"{t1_pre_asm:.2f} s, ILU(0) update {t1_factor:.2f} s, {sample_iters} "
                                  f"iterations timed at {t1_it:.3f} s each and scaled to the {iters} iterations "
                                  f"of the step, postIterate {t1_post:.2f} s")},
        "sample": (f"the first time step of the same workload, whole, on the C oracle (OpenMP, {threads} host "
                   f"threads; ILU(0) factorization and triangular solves sequential): preIterate+assembly "
                   f"{t_pre_asm:.2f} s, ILU(0) update and {iters} FP64 BiCGStab+ILU(0) iterations to 1e-10 "
                   f"{t_solve:.2f} s, postIterate {t_post:.2f} s"),
    }
    return rec, z


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=64, help="cubes per edge")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--delta", type=float, default=0.03, help="FP32 residual drop per FP64 replacement")
    ap.add_argument("--train-steps", type=int, default=10)
    ap.add_argument("--history-depth", type=int, default=16, help="previous increments in the learned guess")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default rows per work-group")
    ap.add_argument("--ilu-steps", type=int, default=4)
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-ilu", action="store_true")
    ap.add_argument("--natural-order", action="store_true",
                    help="iterations in the natural face order (default: faces grouped by type, face_groups = 1)")
    ap.add_argument("--amg-steps", type=int, default=3,
                    help="timed steps of the FP64 BiCGStab + AMG form (0 skips it)")
    ap.add_argument("--isai-steps", type=int, default=3,
                    help="timed steps of the FP64 BiCGStab + ILU-ISAI form (0 skips it)")
    ap.add_argument("--skip-ablation", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=10, help="0 skips the ensemble leg")
    ap.add_argument("--study96", action="store_true",
                    help="only run the large learned-guess study (96^3, time-dependent boundary value)")
    ap.add_argument("--study-size", type=int, default=96)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 1 GiB device buffer 3 times (reference reading for the HBM counters)")
    args = ap.parse_args()

    if args.study96:
        print(json.dumps({"study": study96(args.study_size, [11, 12], 10, 13, 30, args.history_depth)}))
        return 0

    if args.copy_calibration:
        import torch

        n = 1 << 28  # 2^28 floats = 1 GiB
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))
        return 0

    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; a run on several GPUs is not built (it would "
                                   "partition the cells with one ghost layer and exchange face halos and "
                                   "scalar reductions between ranks)", "n_gpus": args.gpus}))
        return 2

    import numpy as np
    from mhfem_amd import capi, problems

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    log(f"[bench] building the n=2 {args.size}^3 workload")
    pr = problems.adr_3d(s=args.size, seed=args.seed, preconditioner="diagonal")
    rows = pr.num_faces * pr.n
    groups = 0 if args.natural_order else 1
    history = dict(train_steps=0, combine="history", history_depth=args.history_depth)
    network = dict(train_steps=args.train_steps, epochs=10, learning_rate=2e-3, sample_stride=8, combine="combined",
                   reject_after=3, history_depth=args.history_depth)
    fp32 = dict(preconditioner="diagonal", precision="fp32-reliable", inner_tol=args.delta, check_every=8, tol=1e-10,
                face_groups=groups)
    chosen = {"name": "previous-increment projection (least residual over the last 16 increments) + FP32 "
                      "BiCGStab with FP64 reliable updates (Jacobi)",
              "solver": fp32, "guess": history}
    with_net = {"name": "the chosen form with the learned network's direction added (trained online on the "
                        "first steps, switched off by its rejection rule)",
                "solver": fp32, "guess": network}
    plain = {"name": "FP64 BiCGStab + Jacobi",
             "solver": dict(preconditioner="diagonal", precision="fp64", check_every=8, tol=1e-10,
                            face_groups=groups)}
    plain_ilu = {"name": "FP64 BiCGStab + ILU(0)",
                 "solver": dict(preconditioner="ilu0", precision="fp64", check_every=4, tol=1e-10)}
    cfg2 = {"name": "FP32 BiCGStab + ILU(0) (FP32 factors) inside FP64 refinement with reliable updates",
            "solver": dict(preconditioner="ilu0", precision="fp32-reliable", inner_tol=args.delta, check_every=4,
                           tol=1e-10)}
    plain_amg = {"name": "FP64 BiCGStab + AMG (PGM aggregation, block-Jacobi Richardson smoothing, V-cycle)",
                 "solver": dict(preconditioner="amg", precision="fp64", check_every=4, tol=1e-10)}
    plain_isai = {"name": "FP64 BiCGStab + ILU-ISAI (ILU(0) factors as sparse approximate inverses, power 2)",
                  "solver": dict(preconditioner="ilu-isai", precision="fp64", check_every=4, tol=1e-10)}
    fp64_hist = {"name": "FP64 BiCGStab + Jacobi from the previous increments",
                 "solver": plain["solver"], "guess": history}
    for f in (chosen, with_net, plain, fp64_hist):
        f["autotune"] = not args.no_autotune

    log(f"[bench] chosen form: {chosen['name']}")
    s, dt, its = run_form(pr, chosen, args.steps, args.warmup, timing=2)
    tm = s.timers()
    spmv_ms = tm["spmv_ms"] / max(1, tm["spmv_timed"])
    rp, col, _, _ = s.export_system()
    nnz = int(rp[-1])
    # stored off-diagonal blocks per face: the distinct other faces in each face's rows
    faces = pr.num_faces
    fc = col.astype(np.int64) // pr.n
    rowface = np.repeat(np.arange(rows, dtype=np.int64) // pr.n, np.diff(rp))
    pairs = np.unique(rowface * faces + fc)
    blocks_per_face = (pairs.size - faces) / faces
    z_chosen = s.get_array("Z_iF").copy()
    s.close()
    res_chosen = summary(chosen["name"], args.steps, dt, its, chosen.get("rows_per_block"), chosen["fallback_steps"])
    log(f"[bench] chosen: {res_chosen}; SpMV {spmv_ms*1e3:.1f} us over {tm['spmv_timed']} timed launches")

    res_plain = z_plain = None
    spmv64_ms = None
    if not args.skip_plain:
        log(f"[bench] plain form: {plain['name']}")
        sp, dtp, itsp = run_form(pr, plain, args.steps, args.warmup, timing=2)
        tmp = sp.timers()
        spmv64_ms = tmp["spmv_ms"] / max(1, tmp["spmv_timed"])
        z_plain = sp.get_array("Z_iF").copy()
        sp.close()
        res_plain = summary(plain["name"], args.steps, dtp, itsp, plain.get("rows_per_block"))
        log(f"[bench] plain: {res_plain}; FP64 SpMV {spmv64_ms*1e3:.1f} us over {tmp['spmv_timed']} timed launches")

    def ilu_leg(problem, form, label):
        log(f"[bench] {label} ({args.ilu_steps} timed steps after 1 warmup step)")
        si, dti, itsi = run_form(problem, form, args.ilu_steps, 1)
        si.close()
        res = summary(form["name"], args.ilu_steps, dti, itsi, None, form["fallback_steps"])
        res["sample"] = f"steps 2-{1 + args.ilu_steps} of the trajectory"
        log(f"[bench] {label}: {res}")
        return res

    res_ilu = res_cfg2 = res_cfg1 = None
    if not args.skip_ilu:
        res_ilu = ilu_leg(pr, plain_ilu, "plain FP64 ILU(0) form (configurations 1-2's preconditioner, n = 2)")
        res_cfg2 = ilu_leg(pr, cfg2, "configuration 2 as configured: FP32 ILU(0) inside FP64 refinement")
        log(f"[bench] building the n=1 Darcy {args.size}^3 workload (configuration 1)")
        pr1 = problems.darcy_3d(s=args.size, seed=args.seed, preconditioner="ilu0")
        c1_ilu = dict(plain_ilu, name="configuration 1: n = 1 Darcy, FP64 BiCGStab + ILU(0)")
        res_cfg1 = ilu_leg(pr1, c1_ilu, "configuration 1 (n = 1 Darcy, FP64 BiCGStab + ILU(0))")
        c1_jac = {"name": "configuration 1 with FP64 BiCGStab + Jacobi",
                  "solver": dict(plain["solver"]), "autotune": not args.no_autotune}
        sj, dtj, itsj = run_form(pr1, c1_jac, args.ilu_steps, 1)
        sj.close()
        res_cfg1["jacobi"] = summary(c1_jac["name"], args.ilu_steps, dtj, itsj, c1_jac.get("rows_per_block"))
        res_cfg1["cells"], res_cfg1["faces"], res_cfg1["dofs"] = pr1.num_cells, pr1.num_faces, pr1.num_faces
        del pr1

    def timed_precond_leg(form, steps, label, extra):
        log(f"[bench] {label} ({steps} timed steps after 1 warmup step, which builds it)")
        sa = capi.Solver(pr, **form["solver"])
        t_setup = time.perf_counter()
        sa.step(0.0, pr.dt)
        sa.synchronize()
        first_ms = 1e3 * (time.perf_counter() - t_setup)
        t = pr.dt
        its0 = len(sa.iterations)
        t0 = time.perf_counter()
        for _ in range(steps):
            sa.step(t, pr.dt)
            t += pr.dt
        sa.synchronize()
        dta = time.perf_counter() - t0
        res = summary(form["name"], steps, dta, sa.iterations[its0:])
        res["first_step_ms_with_setup"] = round(first_ms, 1)
        res.update(extra(sa))
        sa.close()
        log(f"[bench] {label}: {res}")
        return res

    res_amg = res_isai = None
    if args.amg_steps > 0:
        def amg_extra(sa):
            levels, rebuilds = sa.amg_hierarchy()
            return dict(levels=len(levels), level_rows=[lv[0] for lv in levels[:6]], rebuilds=rebuilds)
        res_amg = timed_precond_leg(plain_amg, args.amg_steps, "AMG form", amg_extra)
    if args.isai_steps > 0:
        def isai_extra(sa):
            (lrp, _, _), (urp, _, _), rebuilds = sa.isai_matrices()
            return dict(nnz_lower=int(lrp[-1]), nnz_upper=int(urp[-1]), rebuilds=rebuilds)
        res_isai = timed_precond_leg(plain_isai, args.isai_steps, "ILU-ISAI form", isai_extra)

    res_net = res_h64 = z_h64 = None
    net_stats = None
    if not args.skip_ablation:
        log(f"[bench] ablation: {with_net['name']}")
        sn, dtn, itsn = run_form(pr, with_net, args.steps, args.warmup)
        net_stats = sn.guess_stats()
        sn.close()
        res_net = summary(with_net["name"], args.steps, dtn, itsn, with_net.get("rows_per_block"),
                          with_net["fallback_steps"])
        log(f"[bench] with network: {res_net}; guess {net_stats}")
        log(f"[bench] {fp64_hist['name']}")
        sh, dth, itsh = run_form(pr, fp64_hist, args.steps, args.warmup)
        z_h64 = sh.get_array("Z_iF").copy()
        sh.close()
        res_h64 = summary(fp64_hist["name"], args.steps, dth, itsh, fp64_hist.get("rows_per_block"))
        log(f"[bench] fp64 history: {res_h64}")

    value = args.steps / dt
    csr = spmv_bytes(rows, nnz, True)
    # the timed launches are the iteration's SpMV kernels with their fused
    # dot products, sampled alternately from its two products
    fmt = block_sell_bytes(pr.n, faces, blocks_per_face, True, dot_operands=1.0)
    fmt64 = block_sell_bytes(pr.n, faces, blocks_per_face, False, dot_operands=0.5)
    # priced on the bytes the block-SELL kernel must move; the CSR count of
    # the same product is reported beside it
    achieved = fmt / (spmv_ms * 1e-3) / 1e9
    asm_alg = assembly_bytes(pr.n, pr.fpc, pr.num_cells, faces, nnz)
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "time steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 Krylov iteration / f64 residual, correction and scheme",
        "data": "synthetic (seeded Gaussian ln K field on a structured Kuhn mesh)",
        "config": {
            "workload": f"n=2 advection-diffusion-reaction, {args.size}^3 cubes x 6 Kuhn tets, dt=1e-3, "
                        f"{args.warmup + args.steps}-step trajectory, 1e-10 rel. residual",
            "cells": pr.num_cells, "faces": faces, "dofs": rows, "nnz": nnz,
            "form": chosen["name"], "face_groups": groups,
        },
        "iterations_per_step": res_chosen["iterations_per_step"],
        "dof_per_s": round(rows * value, 1),
        "plain": res_plain,
        "plain_ilu0": res_ilu,
        "config2_as_configured": res_cfg2,
        "config1": res_cfg1,
        "plain_amg": res_amg,
        "plain_ilu_isai": res_isai,
        "with_network": res_net,
        "fp64_history": res_h64,
        "chosen_vs_plain": round(value / res_plain["steps_per_s"], 3) if res_plain else None,
        "chosen_vs_plain_ilu0": round(value / res_ilu["steps_per_s"], 3) if res_ilu else None,
        "chosen_vs_fp64_rel_l2": rel(z_chosen, z_plain) if z_plain is not None else None,
        "fp64_history_vs_fp64_rel_l2": (rel(z_h64, z_plain) if z_plain is not None and z_h64 is not None
                                        else None),
        "network": None if net_stats is None else {
            "network_coefficient_at_end": round(net_stats["coef_network"], 4),
            "trained_steps": net_stats["trained_steps"], "network_active_at_end": net_stats["network_active"],
            "network_gain_at_end": round(net_stats["network_gain"], 5),
            "with_network_vs_chosen": round(res_net["steps_per_s"] / res_chosen["steps_per_s"], 3)},
        "roofline": {
            "kernel": f"k_sspmv<float, {pr.n}, ...> (block-SELL SpMV of the FP32 iteration)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic() if pr.n == 2 and args.size == 64 else None,
            "avg_launch_us": round(spmv_ms * 1e3, 2),
            "timed_launches": tm["spmv_timed"],
            "algorithmic_bytes": int(fmt),
            "csr_equivalent": {"bytes": csr, "GBs": round(csr / (spmv_ms * 1e-3) / 1e9, 1),
                               "note": "the same product priced as an int32-index CSR SpMV "
                                       "(8 nnz + 4 (rows+1) + 8 rows): bytes this format does not move"},
        },
        "plain_spmv_fp64": None if spmv64_ms is None else {
            "kernel": f"k_sspmv<double, {pr.n}, ...> (block-SELL SpMV of the plain FP64 iteration)",
            "avg_launch_us": round(spmv64_ms * 1e3, 2),
            "algorithmic_bytes": int(fmt64),
            "achieved": round(fmt64 / (spmv64_ms * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(fmt64 / (spmv64_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "assembly": {
            "kernels": "k_assemble (interior and boundary rows, assembleLinearSystem)",
            "bound": "hbm",
            "avg_ms": round(tm["assemble"] / args.steps, 4),
            "algorithmic_bytes": asm_alg,
            "achieved": round(asm_alg / (tm["assemble"] / args.steps * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(asm_alg / (tm["assemble"] / args.steps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "pre_iterate_ms": round(tm["pre_iterate"] / args.steps, 4),
            "post_iterate_ms": round(tm["post_iterate"] / args.steps, 4),
        },
        "ensemble": None,
        "cpu_baseline": None,
        "error_vs_oracle_first_step": None,
    }
    if args.ensemble_steps > 0:
        log("[bench] batched ensemble")
        out["ensemble"] = ensemble_leg(args.ensemble_steps, 2)
        log(f"[bench] ensemble: {out['ensemble']}")
    if not args.skip_cpu:
        log("[bench] cpu baseline: the oracle's first step (FP64 BiCGStab + ILU(0) to 1e-10)")
        out["cpu_baseline"], z_oracle = cpu_baseline(problems.adr_3d(s=args.size, seed=args.seed,
                                                                     preconditioner="ilu0"), args.cpu_iters)
        errs = {"note": "relative L2 of Z_iF after the first time step against the all-FP64 oracle's, both "
                        "solved to 1e-10 (the difference of two solutions that meet that tolerance)"}
        for name, form in (("chosen", chosen), ("plain", plain), ("plain_ilu0", plain_ilu),
                           ("config2_as_configured", cfg2), ("with_network", with_net)):
            if "z_first_step" in form:
                errs[name] = rel(form["z_first_step"], z_oracle)
        out["error_vs_oracle_first_step"] = errs
        log(f"[bench] cpu baseline: {out['cpu_baseline']}; errors {errs}")
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
