#!/usr/bin/env python
"""Throughput of the MI355X ERT forward-and-sensitivity path on one GPU.

Workload (default): the batched 2.5D time-lapse configuration -- a ~20k-node graded P1
triangle mesh, 96 surface electrodes at 1 m, ~3000 Wenner-Schlumberger quadrupoles, the
reference wavenumber rule (19 wavenumbers for this survey), and a seeded 365-step
log-conductivity series.  One timed step solves every right-hand side (wavenumber x source
electrode) of ``--batch-steps`` consecutive time steps to a relative residual of 1e-10 and
reduces them to apparent resistivities.  The main solver (``--preconditioner``, default
multigrid-preconditioned flexible CG) and plain Jacobi PCG are timed on the same batches.
After the timed steps, a second timed phase materializes the full normal-sensitivity
Jacobian (all quadrupole rows x all cells) of one time step with the fused kernel.

Prints ONE JSON line.  ``value`` = RHS solves/s of the main solver over the timed steps (wall
clock between two synchronizations, graph-replayed iterations); ``solver.rhs_solves_per_s_pcg``
is plain Jacobi PCG; ``jacobian_rows_per_s`` is the second phase.  Kernel rooflines come from
one more step of the same workload run with the solver's profiling mode: the same kernels
launched directly on the same stream with HIP timing events around every multigrid cycle,
spmv and update launch, so each kernel's average duration is measured live.
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import platform
import sys
import time
from pathlib import Path

os.environ.setdefault("ADTLERT_ENABLE_FLOAT64", "1")
ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(case, series, wavenumbers, weights, sample_steps: int, rows: int):
    """FP64 CPU oracle (scipy splu direct solves, the restatement of the reference path) on a
    bounded sample: ``sample_steps`` time steps (all wavenumbers x sources) and ``rows`` Jacobian rows."""

    import numpy as np

    from oracle import ert2p5d as oracle2p5d

    f = oracle2p5d.Forward2p5D(case.nodes, case.cells, case.electrodes, case.measurements, None, wavenumbers, weights)
    t0 = time.perf_counter()
    for t in range(sample_steps):
        f.fields(series[t])
    solve_s = time.perf_counter() - t0
    rhs = sample_steps * f.W * f.E
    f.slots = f.slots[:rows]
    f.D = rows
    t0 = time.perf_counter()
    f.jacobian(series[sample_steps - 1])
    jac_s = time.perf_counter() - t0
    return rhs / solve_s, rows / jac_s, solve_s, jac_s


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch-steps", type=int, default=4, help="time steps per timed step")
    ap.add_argument("--nodes", type=int, default=20000)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--check-every", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=2, help="time steps in the CPU-baseline sample (0: skip)")
    ap.add_argument("--cpu-rows", type=int, default=32, help="Jacobian rows in the CPU-baseline sample")
    ap.add_argument("--no-jacobian", action="store_true")
    ap.add_argument("--rows-per-tile", type=int, default=None, help="PCG workgroup tile height (default: auto)")
    ap.add_argument("--preconditioner", choices=("amg", "jacobi"), default="amg")
    ap.add_argument("--pcg-steps", type=int, default=2, help="timed steps of plain Jacobi PCG on the same batches (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path shards by wavenumber/time step with one tiny all-reduce; see the design notes)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    adt = importlib.import_module("ad-tlert_amd")
    dev = adt.native.device()
    cus, arch = adt.native.device_info(0)
    total_models = (max(args.warmup + args.steps, 1 + args.pcg_steps) + 1) * args.batch_steps
    t0 = time.perf_counter()
    case, series = adt.cases.timelapse_ws_20k(steps=max(total_models, 2), node_target=args.nodes)
    mesh = adt.Mesh.from_arrays(case.nodes, case.cells)
    survey = adt.Survey.from_arrays(case.electrodes, case.measurements)
    op = adt.ERTForward2p5D.from_mesh_survey(mesh, survey, solver_tolerance=args.tol)
    op._solver.check_every = args.check_every
    op._solver.rows_per_tile = args.rows_per_tile
    d = op._dev
    W, E, N, D = d.W, d.E, d.N, survey.measurement_count
    C = mesh.cell_count
    log(f"setup {time.perf_counter() - t0:.1f}s: N={N} C={C} nnz={d.nnz} W={W} E={E} D={D} device={arch} CUs={cus}")
    sig = torch.as_tensor(series, dtype=torch.float64, device=dev)

    def step(i: int):
        fields = op._solve_fields(sig[i * args.batch_steps:(i + 1) * args.batch_steps])
        stats = op._solver.last
        pots = op._potentials(torch.stack([f.blocks for f in fields]))
        normal, reciprocal = op._normal_reciprocal(pots)
        rhoa = op._geometric_factors().to(dev, torch.float64) * op._combine(normal, reciprocal)
        return fields, rhoa, stats

    def timed(preconditioner: str, first: int, steps: int, warmup: int):
        """Wall-clock of `steps` batches solved with `preconditioner` (after `warmup` untimed ones)."""

        op._solver.preconditioner = preconditioner
        for i in range(first, first + warmup):
            _, _, stats = step(i)
            log(f"[{preconditioner}] warmup {i}: iters mean {stats.iterations.mean():.0f} max "
                f"{stats.iterations.max()} relres max {stats.relative_residual.max():.2e}")
        iters, worst, fallback = [], 0.0, 0
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for i in range(first + warmup, first + warmup + steps):
            fields, rhoa, stats = step(i)
            iters.append(stats.iterations.mean())
            worst = max(worst, float(stats.relative_residual.max()))
            fallback += stats.fallback_rhs
            log(f"[{preconditioner}] step {i}: iters mean {stats.iterations.mean():.1f} max {stats.iterations.max()} "
                f"relres max {stats.relative_residual.max():.2e} fallback {stats.fallback_rhs}")
        torch.cuda.synchronize()
        return time.perf_counter() - t_start, float(np.mean(iters)), worst, fallback

    # the same batches (steps 0 .. warmup+steps-1) for both solvers
    elapsed, mean_iters, worst_relres, fallback = timed(args.preconditioner, 0, args.steps, args.warmup)
    rhs_total = args.steps * args.batch_steps * W * E
    rhs_per_s = rhs_total / elapsed
    pcg_s, pcg_iters, pcg_relres, _ = timed("jacobi", 0, args.pcg_steps, 1 if args.pcg_steps else 0)
    pcg_per_s = args.pcg_steps * args.batch_steps * W * E / pcg_s if args.pcg_steps else None
    op._solver.preconditioner = args.preconditioner
    plan = next(iter(op._solver._plans.values()))
    levels = [plan.amg_level(l)["rows"] for l in range(plan.amg_levels)]
    # a copy of known size (256 MiB read + 256 MiB written): hardware-counter runs of this command
    # scale the solver kernels' FETCH_SIZE / WRITE_SIZE by what the counters read for it
    calib_src = torch.ones(32 * 2 ** 20, dtype=torch.float64, device=dev)
    calib_dst = torch.empty_like(calib_src)
    calib_dst.copy_(calib_src)
    del calib_src, calib_dst
    # profiled step: direct launches with HIP events around every cycle / spmv / update launch
    plan.profile(True)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    _, _, stats = step(args.warmup + args.steps)
    torch.cuda.synchronize()
    profiled_s = time.perf_counter() - tp
    prof = plan.profile_read()
    plan.profile(False)
    it = stats.iterations
    rhs_iters = int(it.sum())
    padded = np.zeros((it.shape[0], plan.rp), dtype=np.int64)
    padded[:, :E] = it
    group_iters = int(padded.reshape(it.shape[0], -1, plan.lanes).max(axis=2).sum())
    # algorithmic bytes of the two streaming Krylov kernels over the profiled step
    nnz = d.nnz
    matrix_bytes = 12 * nnz + 4 * (N + 1) + 8 * N  # values+indices, indptr, D^-1 (gathered)
    spmv_bytes = rhs_iters * N * 32 + group_iters * matrix_bytes  # read r (z), p_old; write p_new, q
    update_bytes = rhs_iters * N * 48 + group_iters * 8 * N  # read x, p, r, q; write x, r
    spmv_gbs = spmv_bytes / (prof["spmv_ms"] * 1e-3) / 1e9
    update_gbs = update_bytes / (prof["update_ms"] * 1e-3) / 1e9
    dominant = ("cg_spmv", spmv_gbs, spmv_bytes, prof["spmv_ms"]) if prof["spmv_ms"] >= prof["update_ms"] else \
        ("cg_update", update_gbs, update_bytes, prof["update_ms"])
    launches = max(prof["iterations"], 1)
    solver_name = {"amg": "FP64 flexible CG, smoothed-aggregation V(1,1) cycle, graph-replayed",
                   "jacobi": "FP64 Jacobi-PCG, graph-replayed"}[args.preconditioner]
    result = {
        "metric": "RHS solves/s + Jacobian rows/s, learned-FCG vs plain PCG at rel-res 1e-10",
        "value": rhs_per_s,
        "unit": "RHS solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded graded P1 mesh, Wenner-Schlumberger survey, GRF time-lapse conductivity)",
        "config": {"workload": "2.5D time-lapse batch: ~20k-node P1 mesh, 96 electrodes, ~3000 Wenner-Schlumberger rows",
                   "nodes": N, "cells": C, "nnz": nnz,
                   "electrodes": E, "quadrupoles": D, "wavenumbers": W, "wavenumber_rule": "reference (integration.py)",
                   "time_steps_per_step": args.batch_steps, "rhs_per_step": args.batch_steps * W * E,
                   "solver": solver_name, "tolerance": args.tol,
                   "lanes": plan.lanes, "rp": plan.rp, "n_tiles": plan.tiles, "rows_per_tile": plan.rows_per_tile,
                   "lds_staged_rows": plan.lds_rows, "dof_order": os.environ.get("ADT_DOF_ORDER", "hilbert"),
                   "multigrid_levels": levels},
        "solver": {"preconditioner": args.preconditioner, "rhs_solves_per_s": rhs_per_s,
                   "rhs_solves_per_s_pcg": pcg_per_s,
                   "speedup_vs_pcg": rhs_per_s / pcg_per_s if pcg_per_s else None,
                   "mean_iterations": mean_iters, "mean_iterations_pcg": pcg_iters if args.pcg_steps else None,
                   "worst_true_relres": max(worst_relres, pcg_relres), "jacobi_fallback_rhs": fallback,
                   "rhs_solves_per_s_learned_fcg": None},
        "roofline": {"bound": "hbm", "kernel": dominant[0], "achieved": dominant[1], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dominant[1] / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": dominant[2] / launches, "avg_launch_ms": dominant[3] / launches,
                     "launches": launches},
        "kernels": {"cg_spmv": {"gbs": spmv_gbs, "avg_ms": prof["spmv_ms"] / launches},
                    "cg_update": {"gbs": update_gbs, "avg_ms": prof["update_ms"] / launches},
                    "multigrid_cycle": {"avg_ms": prof["cycle_ms"] / launches}},
    }
    result["profiled_step"] = {"seconds": profiled_s, "rhs_solves_per_s": args.batch_steps * W * E / profiled_s,
                               "iterations_profiled": prof["iterations"]}
    log(f"timed: {elapsed:.2f}s, {rhs_per_s:.0f} RHS/s ({args.preconditioner}); plain PCG "
        f"{pcg_per_s or 0:.0f} RHS/s; profiled step {profiled_s:.2f}s; cycle {prof['cycle_ms'] / launches:.3f} ms, "
        f"spmv {spmv_gbs:.0f} GB/s ({prof['spmv_ms'] / launches:.3f} ms), "
        f"update {update_gbs:.0f} GB/s ({prof['update_ms'] / launches:.3f} ms)")

    if not args.no_jacobian:
        fields = op._solve_fields(sig[:1])
        columns = op._columns()
        op._normal_jacobian(fields[0], D, columns, return_device=True)  # warm
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 3
        start.record()
        for _ in range(reps):
            jac = op._normal_jacobian(fields[0], D, columns, return_device=True)
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / reps
        result["jacobian_rows_per_s"] = D / (ms * 1e-3)
        K = 3
        sens_bytes = 8 * D * C + 8 * W * E * N + 8 * 2 * C * K * K + 4 * C * K + 16 * D
        sens_flops = D * C * W * (4 * K * K + 6 * K)
        result["roofline_sensitivity"] = {
            "bound": "vector-fp64", "kernel": "sensitivity_kernel<3,double>", "avg_launch_ms": ms,
            "achieved_gbs": sens_bytes / (ms * 1e-3) / 1e9, "hbm_frac": sens_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "achieved_tflops": sens_flops / (ms * 1e-3) / 1e12, "fp64_vector_peak_tflops": 78.6,
            "flops_per_launch": sens_flops, "bytes_per_launch": sens_bytes}
        log(f"jacobian {D}x{C}: {ms:.1f} ms/launch -> {result['jacobian_rows_per_s']:.0f} rows/s")
        del jac

    if args.cpu_steps > 0:
        wn, wt = op.wavenumbers.double().numpy(), op.weights.double().numpy()
        cpu_rhs, cpu_rows, cpu_solve_s, cpu_jac_s = cpu_baseline(case, series, wn, wt, args.cpu_steps, args.cpu_rows)
        result["cpu_baseline"] = {
            "value": cpu_rhs, "unit": "RHS solves/s", "cores": 1, "kind": "port",
            "sample": f"{args.cpu_steps} time steps x {W} wavenumbers x {E} sources, scipy splu FP64 "
                      f"({cpu_solve_s:.1f}s); Jacobian {args.cpu_rows} rows x {C} cells einsum ({cpu_jac_s:.1f}s)",
            "jacobian_rows_per_s": cpu_rows, "cpu": platform.processor() or platform.machine(),
            "host_cores_available": len(os.sched_getaffinity(0))}
        log(f"cpu baseline: {cpu_rhs:.0f} RHS/s, {cpu_rows:.1f} rows/s")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
