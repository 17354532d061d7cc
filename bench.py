#!/usr/bin/env python
"""Throughput of the MI355X ERT forward-and-sensitivity path on one GPU.

Workloads (``--workload``; by default all three, the time-lapse batch as the top-level result, the
3D box under ``box3d`` and the 3D terrain mesh under ``terrain_p2``):

* ``timelapse``: a ~20k-node graded P1 triangle mesh, 96 surface electrodes at 1 m, ~3000
  Wenner-Schlumberger quadrupoles, the reference wavenumber rule (19 wavenumbers for this
  survey) and a seeded 365-step log-conductivity series.  One timed step solves every
  right-hand side (wavenumber x source electrode) of ``--batch-steps`` consecutive time steps.
* ``box3d``: a 200 m x 200 m x 100 m box of 202 005 nodes / 1.15M P1 tetrahedra refined under an
  8 x 8 electrode grid at 5 m, 2000 random quadrupoles, mixed boundary, analytic half-space
  singularity removal, log-normal conductivity fields.  One timed step solves the 64 source
  right-hand sides of ``--batch-steps`` conductivity models.
* ``terrain_p2``: the same kind of box with a relief of 6 seeded sinusoids (up to 5 m) and
  ten-node (P2) tetrahedra, 138 720 cells / 195 201 DOFs, 64 electrodes, numerical geometric
  factors; timed like ``box3d``; the cycle starts with the interpolation from the vertices.

Every step solves to a relative residual of 1e-10 and reduces the fields to apparent
resistivities.  The main solver (``--preconditioner``, default ``learned``: flexible CG
preconditioned by the multigrid cycle whose finest prolongator is damped by the trained graph
network), the same cycle with constant damping (``amg``) and plain Jacobi PCG are timed on the
same batches.  After the timed steps a second timed phase materializes the full
normal-sensitivity Jacobian (all quadrupole rows x all cells) of one model with the fused kernel.

Prints ONE JSON line.  ``value`` = RHS solves/s of the main solver over the timed steps of the
time-lapse batch (wall clock between two synchronizations, graph-replayed iterations, the
network's inference included); ``solver.rhs_solves_per_s_pcg`` is plain Jacobi PCG and
``solver.rhs_solves_per_s_amg`` the constant-damping cycle; ``jacobian_rows_per_s`` is the
second phase.  Kernel rooflines come from one more step of the same workload run with the
solver's profiling mode: the same kernels launched directly on the same stream with HIP timing
events around every multigrid cycle, spmv and update launch, so each kernel's average duration
is measured live.  ``oracle`` gives the maximum relative error of apparent resistivities and
Jacobian rows against the FP64 CPU oracle; ``cpu_baseline`` times the oracle (threaded sparse
LU per (time step, wavenumber) in 2.5D; threaded block Jacobi-PCG in 3D) on the host's cores.
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
METRIC = "RHS solves/s + Jacobian rows/s, learned-FCG vs plain PCG at rel-res 1e-10"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    """Host cores for the CPU baseline: this process's affinity set, at most 16 (a one-GPU share)."""

    return max(1, min(len(os.sched_getaffinity(0)), 16))


def relmax(actual, expected) -> float:
    import numpy as np

    actual, expected = np.asarray(actual, dtype=float), np.asarray(expected, dtype=float)
    return float(np.abs(actual - expected).max() / max(np.abs(expected).max(), 1e-300))


_ORACLE = {}


def _oracle_worker_init(case, wavenumbers, weights, models) -> None:
    """Per-process oracle of the CPU baseline's worker pool (built once per worker)."""

    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import ert2p5d as oracle2p5d

    # one core per worker: BLAS/OpenMP pools of their own would oversubscribe the host
    _ORACLE["limits"] = threadpool_limits(1)
    f = oracle2p5d.Forward2p5D(case.nodes, case.cells, case.electrodes, case.measurements, None, wavenumbers, weights)
    _ORACLE.update(f=f, models=models, rho=[f.source_resistivities(m) for m in models])


def _oracle_worker_ready(_) -> bool:
    return "f" in _ORACLE


def _oracle_worker_system(task):
    """One (time step, wavenumber) system: sparse LU of ``A_w(sigma_t)``, every source solved by
    the secondary-field split, reduced to the weighted ``(E, E)`` electrode potentials."""

    import numpy as np
    import scipy.sparse.linalg as spla

    t, w = task
    f, rho = _ORACLE["f"], _ORACLE["rho"][t]
    a = f.matrix(_ORACLE["models"][t], w).tocsc()
    x0 = f.primary[w] * rho[:, None]
    rhs = f.b[w] - (a @ x0.T).T
    total = spla.splu(a).solve(np.asfortranarray(rhs.T)).T + x0
    return t, f.q[w] * (total @ f.electrodes.T)


def timelapse_cpu(case, series, wavenumbers, weights, sample_steps: int, rows, workers: int):
    """FP64 CPU oracle on a bounded sample: every (time step, wavenumber) system of
    ``sample_steps`` steps factorized by scipy ``splu`` and solved for all sources (the
    restatement of the reference's direct solve), one task per system on a pool of ``workers``
    processes (each builds the oracle once before the timed region), reduced to apparent
    resistivities; then ``rows`` Jacobian rows of step 0 by the oracle's einsum (one core)."""

    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor

    import numpy as np

    from oracle import ert2p5d as oracle2p5d

    models = np.ascontiguousarray(series[:sample_steps])
    f = oracle2p5d.Forward2p5D(case.nodes, case.cells, case.electrodes, case.measurements, None, wavenumbers,
                               weights)
    tasks = [(t, w) for t in range(sample_steps) for w in range(f.W)]
    potentials = np.zeros((sample_steps, f.E, f.E))
    ctx = multiprocessing.get_context("spawn")  # the parent holds a GPU context: no fork
    with ProcessPoolExecutor(workers, mp_context=ctx, initializer=_oracle_worker_init,
                             initargs=(case, wavenumbers, weights, models)) as pool:
        assert all(pool.map(_oracle_worker_ready, range(4 * workers)))  # every worker has built its oracle
        t0 = time.perf_counter()
        for t, part in pool.map(_oracle_worker_system, tasks):
            potentials[t] += part
        solve_s = time.perf_counter() - t0
    k = f.geometric_factors()
    rhoa = []
    for t in range(sample_steps):
        p = np.zeros((f.E + 1, f.E + 1))
        p[: f.E, : f.E] = potentials[t]
        a, b, m, n = f.slots.T
        normal = (p[a, m] - p[a, n]) - (p[b, m] - p[b, n])
        reciprocal = (p[m, a] - p[m, b]) - (p[n, a] - p[n, b])
        rhoa.append(k * np.sign(normal) * np.sqrt(np.abs(normal * reciprocal)))
    f.fields(series[0])  # the Jacobian's fields (not timed)
    t0 = time.perf_counter()
    jac = f.jacobian(series[0], rows=rows, chunk=16)
    jac_s = time.perf_counter() - t0
    return dict(rhs_per_s=len(tasks) * f.E / solve_s, solve_s=solve_s, rows_per_s=len(rows) / jac_s, jac_s=jac_s,
                rhoa=np.stack(rhoa), jacobian=jac, systems=len(tasks), sources=f.E)


class Timelapse:
    """Batched 2.5D time-lapse workload."""

    name = "2.5D time-lapse batch: ~20k-node P1 mesh, 96 electrodes, ~3000 Wenner-Schlumberger rows"
    data = "synthetic (seeded graded P1 mesh, Wenner-Schlumberger survey, GRF time-lapse conductivity)"

    def __init__(self, adt, torch, args, models: int):
        self.args = args
        self.torch = torch
        self.case, self.series = adt.cases.timelapse_ws_20k(steps=max(models, 2), node_target=args.nodes)
        mesh = adt.Mesh.from_arrays(self.case.nodes, self.case.cells)
        survey = adt.Survey.from_arrays(self.case.electrodes, self.case.measurements)
        self.op = adt.ERTForward2p5D.from_mesh_survey(mesh, survey, solver_tolerance=args.tol,
                                                      preconditioner=args.preconditioner)
        d = self.op._dev
        self.W, self.E, self.N, self.nnz = d.W, d.E, d.N, d.nnz
        self.D, self.C = survey.measurement_count, mesh.cell_count
        self.rhs_per_step = args.batch_steps * self.W * self.E
        self.sig = torch.as_tensor(self.series, dtype=torch.float64, device=adt.native.device())
        self.config = {"nodes": self.N, "cells": self.C, "nnz": self.nnz, "electrodes": self.E,
                       "quadrupoles": self.D, "wavenumbers": self.W, "wavenumber_rule": "reference (integration.py)",
                       "time_steps_per_step": args.batch_steps, "rhs_per_step": self.rhs_per_step}

    def step(self, i: int):
        op, n = self.op, self.args.batch_steps
        fields = op._solve_fields(self.sig[i * n:(i + 1) * n])
        stats = op._solver.last
        pots = op._potentials(self.torch.stack([f.blocks for f in fields]))
        normal, reciprocal = op._normal_reciprocal(pots)
        op._geometric_factors().to(self.sig.device, self.torch.float64) * op._combine(normal, reciprocal)
        return stats

    def sensitivity_config(self):
        return {"measurements": self.D, "cells": self.C, "sources": self.E, "wavenumbers": self.W}

    def jacobian(self):
        op = self.op
        fields = op._solve_fields(self.sig[:1])
        columns = op._columns()

        def run():
            return op._normal_jacobian(fields[0], self.D, columns, return_device=True)

        # Gram form: per cell the upper triangle of G_c = Y^T Y over W*K rows of E sources (FP64
        # matrix cores), Y = R Phi with T = R^T R, then 4 lookups and 3 adds per output (a 16-byte
        # table entry per row)
        K, S, rp = 3, self.E, self.op._dev.rp
        bytes_ = 8 * self.D * self.C + 8 * self.W * self.N * rp + 8 * 2 * self.C * K * K + 4 * self.C * K + 16 * self.D
        flops = self.C * (self.W * K * S * (S + 1) + self.W * K * (K + 1) * S) + 3 * self.D * self.C
        return run, bytes_, flops, "gram_kernel<6, 4, 1> (Gram-form sensitivity)", {"bound": "fp64"}

    def cpu_and_oracle(self, jac_device):
        """CPU baseline (oracle on the host cores) and the GPU's error against it on the same
        models: apparent resistivities of the sampled steps and Jacobian rows of step 0."""

        import numpy as np

        args, op, torch = self.args, self.op, self.torch
        if args.cpu_steps <= 0:
            return None, None
        rng = np.random.default_rng(7)
        rows = np.sort(rng.choice(self.D, min(args.cpu_rows, self.D), replace=False))
        threads = cpu_threads()
        wn, wt = op.wavenumbers.double().numpy(), op.weights.double().numpy()
        cpu = timelapse_cpu(self.case, self.series, wn, wt, args.cpu_steps, rows, threads)
        op._field_cache.clear()
        gpu_rhoa = op.apparent_resistivity_series(torch.as_tensor(self.series[: args.cpu_steps])).double().numpy()
        oracle = {"rhoa_max_rel_err": relmax(gpu_rhoa, cpu["rhoa"]),
                  "rhoa_sample": f"{args.cpu_steps} time steps x {self.D} quadrupoles"}
        if jac_device is not None:
            idx = torch.as_tensor(rows, dtype=torch.long, device=jac_device.device)
            oracle["jacobian_max_rel_err"] = relmax(jac_device[idx].cpu().numpy(), cpu["jacobian"])
            oracle["jacobian_sample"] = f"{len(rows)} seeded rows x {self.C} cells of time step 0"
        baseline = {"value": cpu["rhs_per_s"], "unit": "RHS solves/s", "cores": threads, "kind": "port",
                    "sample": f"{args.cpu_steps} time steps x {self.W} wavenumbers x {self.E} sources: "
                              f"{cpu['systems']} scipy splu FP64 factorizations + solves on a pool of {threads} "
                              f"processes ({cpu['solve_s']:.1f}s); Jacobian {len(rows)} rows x {self.C} cells, "
                              f"oracle einsum on one core ({cpu['jac_s']:.1f}s)",
                    "jacobian_rows_per_s": cpu["rows_per_s"], "jacobian_cores": 1,
                    "cpu": platform.processor() or platform.machine(),
                    "host_cores_available": len(os.sched_getaffinity(0))}
        return baseline, oracle


def mixed_batch(adt, torch, args) -> dict:
    """C2's mixed batch: the 96-electrode Wenner-Schlumberger survey on 8 graded meshes of
    5k-40k nodes, one model each, solved in one plan (``MixedMeshBatch``) against the same
    operators solved one after another; apparent resistivity of the smallest mesh against the
    oracle."""

    import numpy as np

    from oracle import ert2p5d as oracle2p5d

    batch = importlib.import_module("ad-tlert_amd.batch")
    cases = adt.cases.mixed_meshes_ws()
    ops = []
    t0 = time.perf_counter()
    for c in cases:
        mesh = adt.Mesh.from_arrays(c.nodes, c.cells)
        survey = adt.Survey.from_arrays(c.electrodes, c.measurements)
        ops.append(adt.ERTForward2p5D.from_mesh_survey(mesh, survey, solver_tolerance=args.tol,
                                                       preconditioner=args.preconditioner))
    mixed = batch.MixedMeshBatch(ops)
    rng = np.random.default_rng(61)
    models = [[torch.as_tensor(c.conductivity * np.exp(0.1 * rng.standard_normal(c.cells.shape[0])))
               for c in cases] for _ in range(args.mixed_steps + 1)]
    rhoa = mixed.apparent_resistivity_values(models[0])  # plans, hierarchies, tuning
    for op, sig in zip(ops, models[0]):
        op.apparent_resistivity_values(sig)
    setup_s = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(1, args.mixed_steps + 1):
        mixed.apparent_resistivity_values(models[i])
    torch.cuda.synchronize()
    together_s = (time.perf_counter() - t0) / args.mixed_steps
    worst = float(mixed.solver.last.relative_residual.max())
    t0 = time.perf_counter()
    for i in range(1, args.mixed_steps + 1):
        for op, sig in zip(ops, models[i]):
            op.apparent_resistivity_values(sig)
    torch.cuda.synchronize()
    alone_s = (time.perf_counter() - t0) / args.mixed_steps
    c = cases[0]
    ref = oracle2p5d.Forward2p5D(c.nodes, c.cells, c.electrodes, c.measurements)
    err = relmax(rhoa[0].double().numpy(), ref.apparent_resistivity(models[0][0].double().numpy()))
    rhs = mixed.rhs_per_solve
    for op in ops:
        op.close()
    mixed.close()
    log(f"mixed batch of {len(cases)} meshes: {rhs / together_s:.0f} RHS/s together, {rhs / alone_s:.0f} one by one")
    return {"meshes_nodes": [int(c.nodes.shape[0]) for c in cases], "rhs_per_solve": rhs,
            "rhs_solves_per_s": rhs / together_s, "rhs_solves_per_s_one_by_one": rhs / alone_s,
            "speedup_vs_one_by_one": alone_s / together_s, "worst_true_relres": worst,
            "oracle_rhoa_max_rel_err_smallest_mesh": err, "timed_solves": args.mixed_steps,
            "setup_s": setup_s, "preconditioner": args.preconditioner}


class Box3D:
    """3D P1 box workload: batches of conductivity models, 64 source right-hand sides each."""

    name = "3D P1 box: 202 005 nodes / 1.15M tetrahedra, 8x8 electrodes at 5 m, 2000 random quadrupoles"
    data = "synthetic (seeded graded tetrahedral box, random ABMN survey, log-normal GRF conductivity models)"
    order = 1
    fixture = "box_p1_200k"

    @staticmethod
    def make_case(adt):
        return adt.cases.box_p1_200k()

    def __init__(self, adt, torch, args, models: int):
        import numpy as np

        self.args = args
        self.torch = torch
        self.case = self.make_case(adt)
        mesh = adt.Mesh3D.from_arrays(self.case.nodes, self.case.cells)
        survey = adt.Survey.from_arrays(self.case.electrodes, self.case.measurements)
        self.op = adt.ERTForward3D.from_mesh_survey(mesh, survey, element_order=self.order, boundary_mode="mixed",
                                                    solver_tolerance=args.tol, series_batch_steps=args.batch_steps,
                                                    preconditioner=args.preconditioner)
        s = self.op._dev  # builds the device state (pattern, templates, primaries)
        self.N, self.nnz, self.C = s.N, s.nnz, mesh.cell_count
        self.D, self.E = survey.measurement_count, survey.electrode_count
        self.rhs_per_step = args.batch_steps * self.op._dev.nc
        # models: the seeded base field times two smooth log-perturbations with varying weights
        rng = np.random.default_rng(44)
        centers = self.case.nodes[self.case.cells].mean(axis=1)
        g1 = np.log10(adt.cases.grf_3d(centers, rng, mean=0.0, std=1.0))
        g2 = np.log10(adt.cases.grf_3d(centers, rng, mean=0.0, std=1.0))
        t = np.arange(max(models, 2))[:, None]
        logs = np.log10(self.case.conductivity)[None] + 0.2 * (np.sin(0.7 * t) * g1[None] + np.cos(0.3 * t) * g2[None])
        self.series_host = 10.0 ** logs
        self.sig = torch.as_tensor(self.series_host, dtype=torch.float64, device=adt.native.device())
        self.config = {"nodes": self.N, "cells": self.C, "nnz": self.nnz, "electrodes": self.E,
                       "quadrupoles": self.D, "source_rhs_per_model": self.op._dev.nc,
                       "models_per_step": args.batch_steps, "rhs_per_step": self.rhs_per_step,
                       "boundary": "mixed", "singularity_removal": True, "element_order": self.order}

    def step(self, i: int):
        op, n = self.op, self.args.batch_steps
        fields = op._solve_fields(self.sig[i * n:(i + 1) * n])
        stats = op._solver.last
        current = self.torch.stack([f.current for f in fields])
        op._geometric_factors().to(self.sig.device, self.torch.float64) * op._measurement_resistances(current)
        return stats

    def sensitivity_config(self):
        return {"measurements": self.D, "cells": self.C, "receivers": self.op._dev.nr, "currents": self.op._dev.nc}

    def jacobian(self):
        op = self.op
        sigma = self.sig[0].cpu()
        op.jacobian(sigma, return_device=True, batch_size=self.D)  # fields into the cache

        def run():
            return op.jacobian(sigma, return_device=True, batch_size=self.D)

        K, nc, nr = (4 if self.order == 1 else 10), op._dev.nc, op._dev.nr
        bytes_ = 8 * self.D * self.C + 8 * self.N * (nc + nr) + 8 * self.C * K * K + 4 * self.C * K + 16 * self.D
        if self.gram_form:
            # Gram form: per cell Psi = T Phi_C (K x nc) and G_c = Phi_R^T Psi ((nr+1) x (nc+1), on
            # the matrix cores), then per output four LDS reads and three adds; the (D, C) block is
            # written one 8-byte column entry per row
            flops = self.C * (2 * K * K * nc + 2 * K * nr * nc) + 3 * self.D * self.C
            lds_bytes = 4 * 8 * self.D * self.C
        else:
            # cell-block form: per cell Psi = T Phi_C (K x nc), then per output a K-term sum of
            # (receiver difference) x (Psi difference), four LDS reads per term; the 16 GB output
            # writes bound it (a variant with the stores alone takes ~80% of its time)
            flops = self.C * 2 * K * K * nc + 4 * K * self.D * self.C
            lds_bytes = 4 * K * 8 * self.D * self.C
        return run, bytes_, flops, self.sensitivity_kernel, {"bound": "hbm", "lds_bytes_per_launch": lds_bytes}

    sensitivity_kernel = "cell_block_kernel<32, 1024, 4> (cell-block sensitivity)"
    gram_form = False

    def cpu_and_oracle(self, jac_device):
        """CPU baseline: the oracle's matrix solved for ``--cpu-rhs-3d`` of the 64 source
        right-hand sides of model 0 by block Jacobi-PCG on the host cores (the oracle's direct
        factorization takes minutes on this mesh).  Oracle errors: the GPU's apparent
        resistivities and sampled Jacobian rows of the seeded model against the committed
        fixture the oracle's direct solves wrote (``tests/golden/box_p1_200k.npz``)."""

        import numpy as np

        from oracle import cpu_pcg, ert3d

        args, op, torch = self.args, self.op, self.torch
        oracle = None
        fixture = ROOT / "tests" / "golden" / f"{self.fixture}.npz"
        if fixture.exists():
            data = np.load(fixture, allow_pickle=False)
            sigma = torch.as_tensor(self.case.conductivity)
            rhoa = op.apparent_resistivity_values(sigma).double().numpy()
            jac = op.jacobian(sigma, return_device=True)
            rows = torch.as_tensor(data["jacobian_rows"], dtype=torch.long, device=jac.device)
            cells = torch.as_tensor(data["jacobian_cells"], dtype=torch.long, device=jac.device)
            oracle = {"rhoa_max_rel_err": relmax(rhoa, data["rhoa"]),
                      "rhoa_sample": f"{self.D} quadrupoles of the seeded model (oracle: sparse LU, committed fixture)",
                      "jacobian_max_rel_err": relmax(jac[rows][:, cells].cpu().numpy(), data["jacobian"]),
                      "jacobian_sample": f"{len(data['jacobian_rows'])} rows x {len(data['jacobian_cells'])} cells "
                                         "(the largest entries plus seeded cells)"}
            del jac
        if args.cpu_rhs_3d <= 0:
            return None, oracle
        threads = cpu_threads()
        ref = ert3d.Forward3D(self.case.nodes, self.case.cells, self.case.electrodes, self.case.measurements)
        matrix = ref._assemble_matrix(self.series_host[0]).tocsr()
        rhs = np.ascontiguousarray(ref.source_rhs[: args.cpu_rhs_3d].T)
        t0 = time.perf_counter()
        _, iterations = cpu_pcg.jacobi_pcg(matrix, rhs, tol=args.tol, threads=threads)
        cpu_s = time.perf_counter() - t0
        baseline = {"value": args.cpu_rhs_3d / cpu_s, "unit": "RHS solves/s", "cores": threads, "kind": "port",
                    "sample": f"{args.cpu_rhs_3d} of the 64 source right-hand sides of model 0, block Jacobi-PCG "
                              f"FP64 to {args.tol:g} ({int(iterations.mean())} iterations; torch CPU sparse products "
                              f"on {threads} threads, {cpu_s:.1f}s)",
                    "cpu": platform.processor() or platform.machine(),
                    "host_cores_available": len(os.sched_getaffinity(0))}
        return baseline, oracle


class TerrainP2(Box3D):
    """3D P2 terrain workload (ten-node tetrahedra, sinusoid relief, numerical geometric factors).
    The cycle's finest prolongator is the interpolation from the vertices, so the learned
    damping (which acts on aggregation prolongators) has nothing to damp and the "learned" leg
    runs the constant cycle.  Oracle errors against the committed fixture of the seeded model;
    no CPU baseline (the 3D box carries one)."""

    name = ("3D P2 terrain: 138 720 ten-node tetrahedra / 195 201 DOFs, sinusoid relief up to 5 m, 64 electrodes, "
            "2000 random quadrupoles, numerical geometric factors")
    data = "synthetic (seeded graded terrain box, random ABMN survey, log-normal GRF conductivity models)"
    order = 2
    fixture = "terrain_p2_150k"
    sensitivity_kernel = "gram_cross_kernel<4> (Gram-form cross sensitivity, P2)"
    gram_form = True

    @staticmethod
    def make_case(adt):
        return adt.cases.terrain_p2_150k()

    def cpu_and_oracle(self, jac_device):
        args = self.args
        saved = args.cpu_rhs_3d
        args.cpu_rhs_3d = 0
        try:
            return super().cpu_and_oracle(jac_device)
        finally:
            args.cpu_rhs_3d = saved


WORKLOADS = {"timelapse": "Timelapse", "box3d": "Box3D", "terrain_p2": "TerrainP2"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("all", "timelapse", "box3d", "terrain_p2"), default="all",
                    help="all: the time-lapse batch as the top-level result, the 3D P1 box under 'box3d' and "
                         "the 3D P2 terrain mesh under 'terrain_p2'")
    ap.add_argument("--batch-steps", type=int, default=None,
                    help="time steps (timelapse, default 4) or models (box3d, terrain_p2: default 2) per timed step")
    ap.add_argument("--nodes", type=int, default=20000)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--check-every", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=8,
                    help="time steps in the 2.5D CPU-baseline sample and its oracle comparison (0: skip)")
    ap.add_argument("--cpu-rows", type=int, default=32, help="Jacobian rows in the 2.5D CPU-baseline sample")
    ap.add_argument("--mixed-steps", type=int, default=3,
                    help="timed solves of the 8-mesh mixed batch with the time-lapse workload (0: skip)")
    ap.add_argument("--cpu-rhs-3d", type=int, default=16,
                    help="source right-hand sides in the 3D CPU-baseline sample (0: skip)")
    ap.add_argument("--no-jacobian", action="store_true")
    ap.add_argument("--rows-per-tile", type=int, default=None, help="PCG workgroup tile height (default: auto)")
    ap.add_argument("--lanes", type=int, default=None, help="lane-group width 8/16/32/64 (default: auto)")
    ap.add_argument("--lds-bytes", type=int, default=None, help="LDS budget of the staged kernels (0: gather kernels)")
    ap.add_argument("--preconditioner", choices=("learned", "amg", "jacobi"), default="learned")
    ap.add_argument("--pcg-steps", type=int, default=2, help="timed steps of plain Jacobi PCG on the same batches (0: skip)")
    ap.add_argument("--amg-steps", type=int, default=2,
                    help="timed steps of flexible CG with the constant-damping cycle on the same batches (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path shards by wavenumber/time step with one tiny all-reduce; see the design notes)",
              file=sys.stderr)
        return 2
    import torch

    adt = importlib.import_module("ad-tlert_amd")
    names = list(WORKLOADS) if args.workload == "all" else [args.workload]
    results = []
    for name in names:
        results.append(measure(adt, torch, args, name))
        torch.cuda.empty_cache()
    result = results[0]
    for name, extra in zip(names[1:], results[1:]):
        result[name] = {k: v for k, v in extra.items() if k not in ("n_gpus", "steps", "warmup", "higher_is_better",
                                                                   "scaling", "vs_baseline")}
    print(json.dumps(result), flush=True)
    return 0


def measure(adt, torch, args, workload: str) -> dict:
    """One workload's measured result (the JSON object of the bench line)."""

    import numpy as np

    batch_steps = args.batch_steps or (4 if workload == "timelapse" else 2)
    args = argparse.Namespace(**{**vars(args), "batch_steps": batch_steps, "workload": workload})

    dev = adt.native.device()
    cus, arch = adt.native.device_info(0)
    models = (max(args.warmup + args.steps, 1 + args.pcg_steps, 1 + args.amg_steps) + 1) * args.batch_steps
    t0 = time.perf_counter()
    wl = globals()[WORKLOADS[args.workload]](adt, torch, args, models)
    op = wl.op
    op._solver.check_every = args.check_every
    op._solver.rows_per_tile = args.rows_per_tile
    op._solver.lanes = args.lanes
    op._solver.lds_bytes = args.lds_bytes
    log(f"setup {time.perf_counter() - t0:.1f}s: {args.workload} N={wl.N} C={wl.C} nnz={wl.nnz} D={wl.D} "
        f"rhs/step={wl.rhs_per_step} device={arch} CUs={cus}")

    def timed(preconditioner: str, first: int, steps: int, warmup: int):
        """Wall-clock of `steps` batches solved with `preconditioner` (after `warmup` untimed ones)."""

        op._solver.preconditioner = preconditioner
        for i in range(first, first + warmup):
            stats = wl.step(i)
            log(f"[{preconditioner}] warmup {i}: iters mean {stats.iterations.mean():.0f} max "
                f"{stats.iterations.max()} relres max {stats.relative_residual.max():.2e}")
        iters, worst, fallback, polished = [], 0.0, 0, 0
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for i in range(first + warmup, first + warmup + steps):
            stats = wl.step(i)
            iters.append(stats.iterations.mean())
            worst = max(worst, float(stats.relative_residual.max()))
            fallback += stats.fallback_rhs
            polished += stats.polished_rhs
            log(f"[{preconditioner}] step {i}: iters mean {stats.iterations.mean():.1f} max {stats.iterations.max()} "
                f"relres max {stats.relative_residual.max():.2e} fallback {stats.fallback_rhs} "
                f"polished {stats.polished_rhs}")
        torch.cuda.synchronize()
        return time.perf_counter() - t_start, float(np.mean(iters)) if iters else None, worst, fallback, polished

    # the same batches (steps 0 .. warmup+steps-1) for both solvers
    elapsed, mean_iters, worst_relres, fallback, polished = timed(args.preconditioner, 0, args.steps, args.warmup)
    rhs_per_s = args.steps * wl.rhs_per_step / elapsed
    pcg_per_s = pcg_iters = None
    pcg_relres = pcg_polished = None
    if args.pcg_steps > 0 and args.preconditioner != "jacobi":
        pcg_s, pcg_iters, pcg_relres, _, pcg_polished = timed("jacobi", 0, args.pcg_steps, 1)
        pcg_per_s = args.pcg_steps * wl.rhs_per_step / pcg_s
    elif args.preconditioner == "jacobi":
        pcg_per_s, pcg_iters = rhs_per_s, mean_iters
    amg_per_s = amg_iters = amg_relres = None
    if args.amg_steps > 0 and args.preconditioner == "learned":
        amg_s, amg_iters, amg_relres, _, _ = timed("amg", 0, args.amg_steps, 1)
        amg_per_s = args.amg_steps * wl.rhs_per_step / amg_s
    op._solver.preconditioner = args.preconditioner
    # a copy of known size (256 MiB read + 256 MiB written): hardware-counter runs of this command
    # scale the solver kernels' FETCH_SIZE / WRITE_SIZE by what the counters read for it
    calib_src = torch.ones(32 * 2 ** 20, dtype=torch.float64, device=dev)
    calib_dst = torch.empty_like(calib_src)
    calib_dst.copy_(calib_src)
    del calib_src, calib_dst
    # profiled step: direct launches with HIP events around every cycle / spmv / update launch
    plan = op._solver.last_plan
    plan.profile(True)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    stats = wl.step(args.warmup + args.steps)
    torch.cuda.synchronize()
    profiled_s = time.perf_counter() - tp
    prof = plan.profile_read()
    plan.profile(False)
    levels = [plan.amg_level(lv)["rows"] for lv in range(plan.amg_levels)]
    it = stats.iterations
    rhs_iters = int(it.sum())
    r = it.shape[1]
    padded = np.zeros((it.shape[0], plan.rp), dtype=np.int64)
    padded[:, :r] = it
    group_iters = int(padded.reshape(it.shape[0], -1, plan.lanes).max(axis=2).sum())
    # algorithmic bytes of the two streaming Krylov kernels over the profiled step
    N, nnz = wl.N, wl.nnz
    matrix_bytes = 12 * nnz + 4 * (N + 1) + 8 * N  # values+indices, indptr, D^-1 (gathered)
    spmv_bytes = rhs_iters * N * 32 + group_iters * matrix_bytes  # read r (z), p_old; write p_new, q
    # the multigrid FCG fuses its update with the next cycle's fine pre-smoothing: it also writes
    # the smoothed residual r - A w D^-1 r and reads the fine matrix once per group
    fused = args.preconditioner != "jacobi" and plan.lds_rows > 0 and plan.amg_levels > 0
    if fused:
        update_name = "fcg_update_smooth"
        update_bytes = rhs_iters * N * 56 + group_iters * matrix_bytes  # read x, p, r, q; write x, r, r_s
    else:
        update_name = "cg_update"
        update_bytes = rhs_iters * N * 48 + group_iters * 8 * N  # read x, p, r, q; write x, r
    spmv_gbs = spmv_bytes / (prof["spmv_ms"] * 1e-3) / 1e9
    update_gbs = update_bytes / (prof["update_ms"] * 1e-3) / 1e9
    dominant = ("cg_spmv", spmv_gbs, spmv_bytes, prof["spmv_ms"]) if prof["spmv_ms"] >= prof["update_ms"] else \
        (update_name, update_gbs, update_bytes, prof["update_ms"])
    launches = max(prof["iterations"], 1)
    solver_name = {"learned": "FP64 flexible CG, filtered smoothed-aggregation V(1,1) cycle with the finest "
                              "prolongator damped per slot by the trained graph network, graph-replayed",
                   "amg": "FP64 flexible CG, filtered smoothed-aggregation V(1,1) cycle, graph-replayed",
                   "jacobi": "FP64 Jacobi-PCG, graph-replayed"}[args.preconditioner]
    config = {"workload": wl.name, **wl.config, "solver": solver_name, "tolerance": args.tol,
              "lanes": plan.lanes, "rp": plan.rp, "n_tiles": plan.tiles, "rows_per_tile": plan.rows_per_tile,
              "lds_staged_rows": plan.lds_rows, "dof_order": os.environ.get("ADT_DOF_ORDER", "hilbert"),
              "multigrid_levels": levels,
              "lds_budget": plan.lds_bytes,
              "autotune_ms_per_iteration": [{f"{k[0]} lanes, {k[1] // 1024} KB": round(v * 1e3, 4) for k, v in t.items()}
                                                 for t in op._solver.tuning.values()]}
    result = {
        "metric": METRIC,
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
        "data": wl.data,
        "config": config,
        "solver": {"preconditioner": args.preconditioner, "rhs_solves_per_s": rhs_per_s,
                   "rhs_solves_per_s_pcg": pcg_per_s,
                   "speedup_vs_pcg": rhs_per_s / pcg_per_s if pcg_per_s else None,
                   "mean_iterations": mean_iters, "mean_iterations_pcg": pcg_iters,
                   "worst_true_relres": worst_relres, "worst_true_relres_pcg": pcg_relres,
                   "worst_true_relres_amg": amg_relres, "jacobi_fallback_rhs": fallback,
                   "true_residual_restarts": polished, "true_residual_restarts_pcg": pcg_polished,
                   "rhs_solves_per_s_amg": amg_per_s, "mean_iterations_amg": amg_iters,
                   "speedup_amg_vs_pcg": amg_per_s / pcg_per_s if amg_per_s and pcg_per_s else None},
        "roofline": {"bound": "hbm", "kernel": dominant[0], "achieved": dominant[1], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dominant[1] / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": dominant[2] / launches, "avg_launch_ms": dominant[3] / launches,
                     "launches": launches},
        "kernels": {"cg_spmv": {"gbs": spmv_gbs, "avg_ms": prof["spmv_ms"] / launches},
                    update_name: {"gbs": update_gbs, "avg_ms": prof["update_ms"] / launches},
                    "multigrid_cycle": {"avg_ms": prof["cycle_ms"] / launches}},
        "profiled_step": {"seconds": profiled_s, "rhs_solves_per_s": wl.rhs_per_step / profiled_s,
                          "iterations_profiled": prof["iterations"]},
    }
    # HBM traffic of the dominant kernel from the committed hardware-counter runs of this same
    # configuration (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a copy of known size)
    traffic_file = ROOT / "profiles" / f"{args.workload}_{dominant[0]}_traffic.json"
    if traffic_file.exists():
        t = json.loads(traffic_file.read_text())
        c = t["config"]
        if (c.get("nodes"), c.get("nnz"), c.get("rp"), c.get("batch_steps"), c.get("preconditioner"),
                c.get("check_every")) == (N, nnz, plan.rp, args.batch_steps, args.preconditioner, args.check_every):
            result["roofline"]["traffic"] = t["traffic_bytes_per_dispatch"]
            result["roofline"]["traffic_source"] = f"profiles/{traffic_file.name}"
    log(f"timed: {elapsed:.2f}s, {rhs_per_s:.0f} RHS/s ({args.preconditioner}); plain PCG "
        f"{pcg_per_s or 0:.0f} RHS/s; constant-damping FCG {amg_per_s or 0:.0f} RHS/s; profiled step {profiled_s:.2f}s; cycle {prof['cycle_ms'] / launches:.3f} ms, "
        f"spmv {spmv_gbs:.0f} GB/s ({prof['spmv_ms'] / launches:.3f} ms), "
        f"{update_name} {update_gbs:.0f} GB/s ({prof['update_ms'] / launches:.3f} ms)")

    jac = None
    if not args.no_jacobian:
        run, sens_bytes, sens_flops, kernel, sens_extra = wl.jacobian()
        run()  # warm (the output block then comes from torch's cache in every timed call)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 3
        jac = None
        start.record()
        for _ in range(reps):
            jac = None  # release the previous output first, so no call allocates a second block
            jac = run()
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / reps
        result["jacobian_rows_per_s"] = wl.D / (ms * 1e-3)
        result["roofline_sensitivity"] = {
            "kernel": kernel, "avg_launch_ms": ms,
            "achieved_gbs": sens_bytes / (ms * 1e-3) / 1e9, "hbm_frac": sens_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "achieved_tflops": sens_flops / (ms * 1e-3) / 1e12, "fp64_peak_tflops": 78.6,
            "flops_per_launch": sens_flops, "bytes_per_launch": sens_bytes, **sens_extra}
        if "lds_bytes_per_launch" in sens_extra:
            result["roofline_sensitivity"]["achieved_lds_tbs"] = sens_extra["lds_bytes_per_launch"] / (ms * 1e-3) / 1e12
        # measured HBM traffic of the Jacobian kernel (committed FETCH_SIZE / WRITE_SIZE runs of this
        # same configuration, scaled by a copy of known size)
        sens_file = ROOT / "profiles" / f"{args.workload}_{kernel.split('<')[0]}_traffic.json"
        if sens_file.exists():
            t = json.loads(sens_file.read_text())
            if t["kernel"] == kernel.split(" (")[0] and t["config"] == wl.sensitivity_config():
                result["roofline_sensitivity"]["traffic"] = t["traffic_bytes_per_dispatch"]
                result["roofline_sensitivity"]["traffic_source"] = f"profiles/{sens_file.name}"
        log(f"jacobian {wl.D}x{wl.C}: {ms:.1f} ms/launch -> {result['jacobian_rows_per_s']:.0f} rows/s")

    if workload == "timelapse" and args.mixed_steps > 0:
        result["mixed_batch"] = mixed_batch(adt, torch, args)
    jac_device = jac if workload == "timelapse" else None
    jac = None  # the 3D comparison computes its own block; release this one first
    baseline, oracle = wl.cpu_and_oracle(jac_device)
    jac_device = None
    if oracle is not None:
        result["oracle"] = oracle
        log(f"oracle: rhoa max rel err {oracle['rhoa_max_rel_err']:.2e}, "
            f"jacobian max rel err {oracle.get('jacobian_max_rel_err', float('nan')):.2e}")
    if baseline is not None:
        result["cpu_baseline"] = baseline
        log(f"cpu baseline: {baseline['value']:.1f} RHS/s on {baseline['cores']} cores")
    op.close()
    return result


if __name__ == "__main__":
    sys.exit(main())
