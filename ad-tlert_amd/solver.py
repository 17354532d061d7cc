"""Batched iterative SPD solves on the GPU (replaces the cuDSS ``BatchedSolver``).

The reference factorizes every matrix with cuDSS and back-substitutes all right-hand sides
(``forward/cudss.py:63-122``).  Here each system is solved in the native library by FP64
conjugate gradients: Jacobi-preconditioned (``preconditioner="jacobi"``), or flexible CG
preconditioned by one smoothed-aggregation multigrid V-cycle per iteration (``"amg"``), whose
hierarchy is built once per plan from the first matrices solved with it and whose numeric
operators are recomputed for every system of every solve; right-hand sides that stall under
the cycle are finished by Jacobi PCG.  A plan per batch layout (number of systems, rows,
right-hand sides) owns the Krylov workspace and the captured iteration graphs and is reused.

Vectors are interleaved ``(N, Rp)`` blocks (node-major, sources contiguous).  ``Rp`` pads the
source count to a multiple of the lane-group width ``L`` chosen by :func:`choose_lanes`;
padding columns are zero and are never iterated.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import native

LANE_CHOICES = (64, 32, 16, 8)


def choose_lanes(r: int) -> int:
    """Lane-group width minimising the padded source count (ties: wider groups)."""

    best = None
    for lanes in LANE_CHOICES:
        rp = -(-r // lanes) * lanes
        if best is None or rp < best[0]:
            best = (rp, lanes)
    return best[1]


def padded(r: int, lanes: int) -> int:
    return -(-r // lanes) * lanes


def choose_rows_per_tile(n_sys: int, rows: int, r: int, lanes: int, target_tiles: int = 16384) -> int:
    """Rows per workgroup tile: at least ``target_tiles`` tiles over the batch (many more than
    the 2048 resident workgroups, so the last round of tiles is short), at least one pass of the
    256-thread block and at most 256 rows (the LDS-staged product may shrink it further)."""

    groups = padded(r, lanes) // lanes
    per_pass = 256 // lanes
    rb = -(-(n_sys * groups * rows) // target_tiles)
    rb = max(per_pass, min(256, rb))
    return -(-rb // per_pass) * per_pass


PRECONDITIONERS = ("jacobi", "amg", "learned", "ic0")


@dataclass
class SolveStats:
    iterations: np.ndarray
    relative_residual: np.ndarray
    seconds: float
    preconditioner: str = "jacobi"
    fallback_rhs: int = 0  # finished by Jacobi PCG after the cycle stalled
    polished_rhs: int = 0  # restarted once because the true residual missed the tolerance


@dataclass
class BatchedSolver:
    """PCG plans keyed by ``(key, n_sys, rows, r)``; ``close()`` releases their workspaces."""

    tolerance: float = 1e-10
    max_iterations: int = 50000
    check_every: int = 4  # iterations per graph replay; the host polls one replay behind
    rows_per_tile: int | None = None
    lanes: int | None = None       # lane-group width (None: choose_lanes)
    lds_bytes: int | None = None   # LDS budget of the staged products (None: native.LDS_BYTES; 0: gather kernels)
    preconditioner: str = "jacobi"
    amg_theta: float = 0.08
    amg_coarse_max: int = 128      # coarsest operators up to 128 rows are inverted in LDS
    amg_stall_limit: int = 64
    amg_smoother_scale: float = native.SMOOTHER_SCALE  # Jacobi smoother damping / (4 / (3 lambda_max))
    damping_network: object = None  # learned.DampingNetwork for preconditioner="learned"
    # a given finest-level prolongator (indptr, indices, values, ncols) for plans of its row count
    # whose systems share one CSR pattern, e.g. P2 -> P1 interpolation; coarser levels aggregate
    first_prolongator: tuple | None = None
    ic0_factor_sweeps: int = 3     # fixed-point sweeps of the IC(0) factor
    ic0_solve_sweeps: int = 2      # Jacobi sweeps per triangular solve of the IC(0) preconditioner
    autotune: bool = True          # time the candidate lane widths before the first plan of a shape
    tuning: dict = field(default_factory=dict)  # shape -> {lanes: seconds per iteration}
    _plans: dict = field(default_factory=dict)
    last: SolveStats | None = None
    last_plan: native.PcgPlan | None = None

    def __post_init__(self):
        if self.preconditioner not in PRECONDITIONERS:
            raise ValueError(f"preconditioner must be one of {PRECONDITIONERS}, got {self.preconditioner!r}")

    def plan(self, key, rows, r: int, *, pattern_offsets=None, b_off=None, lanes=None, rows_per_tile=None,
             host_pattern=None):
        """Plan for systems with ``rows[s]`` rows sharing one CSR pattern unless ``pattern_offsets``
        ``(ptr_off, idx_off, val_off)`` says otherwise."""

        rows = np.asarray(rows, dtype=np.int32)
        n_sys = int(rows.size)
        if lanes is None:
            lanes = self.lanes
        lds = native.LDS_BYTES if self.lds_bytes is None else int(self.lds_bytes)
        if lanes is None:
            lanes, lds = self._tuned_shape(rows, r, pattern_offsets, host_pattern, lds)
        lanes = int(lanes)
        if rows_per_tile is None:
            rows_per_tile = self.rows_per_tile
        if rows_per_tile is None:
            rows_per_tile = choose_rows_per_tile(n_sys, int(rows.max()), r, lanes)
        bkey = None if b_off is None else tuple(np.asarray(b_off).tolist())
        full_key = (key, tuple(rows.tolist()), int(r), lanes, int(rows_per_tile), bkey, lds)
        plan = self._plans.get(full_key)
        if plan is None:
            if pattern_offsets is None:
                ptr_off = np.zeros(n_sys, dtype=np.int64)
                idx_off = np.zeros(n_sys, dtype=np.int64)
                val_off = None
            else:
                ptr_off, idx_off, val_off = pattern_offsets
            if val_off is None:
                raise ValueError("val_off is required")
            plan = native.PcgPlan(rows, r, lanes, rows_per_tile, ptr_off, idx_off, val_off, b_off, host_pattern,
                                  lds_bytes=lds)
            self._plans[full_key] = plan
        return plan

    def _tuned_shape(self, rows, r: int, pattern_offsets, host_pattern, lds: int) -> tuple[int, int]:
        """Lane-group width and LDS budget of the staged kernels for a new problem shape.  Every
        width that keeps the padded source count of :func:`choose_lanes` is timed on the plan's
        CSR pattern (graph-Laplacian values, random right-hand sides, the solver's
        preconditioner, a fixed number of iterations), at the default budget and -- when that
        budget cut the tile height (dense rows: 3D, quadratic elements) and no budget was set --
        at the wide budget ``native.LDS_BYTES_WIDE`` too; the fastest per iteration wins (one
        tuning per shape, whatever the preconditioner).  Narrow groups and wide budgets stage
        taller tiles (fewer halo rows per row) with fewer workgroups per CU; wide groups read the
        matrix fewer times."""

        default = choose_lanes(r)
        rp = padded(r, default)
        candidates = [L for L in LANE_CHOICES if L >= 16 and padded(r, L) == rp]
        if (not self.autotune or host_pattern is None or pattern_offsets is None or len(candidates) < 2
                or not torch.cuda.is_available()):
            return default, lds
        budgets = [lds] if self.lds_bytes is not None or lds <= 0 else [lds, max(lds, native.LDS_BYTES_WIDE)]
        ptr_off, idx_off, _ = (np.asarray(o, dtype=np.int64) if o is not None else None for o in pattern_offsets)
        shape = (tuple(rows.tolist()), int(r), tuple(ptr_off.tolist()), tuple(idx_off.tolist()), lds,
                 self.rows_per_tile)
        if shape in self.tuning:
            times = self.tuning[shape]
            return min(times, key=times.get)
        indptr_h, indices_h = (np.ascontiguousarray(a, dtype=np.int32) for a in host_pattern)
        values, val_off = _laplacian(indptr_h, indices_h, rows, ptr_off, idx_off)
        device = native.device()
        dev_ptr, dev_idx = (torch.as_tensor(a, device=device) for a in (indptr_h, indices_h))
        values_d = torch.as_tensor(values, device=device)
        times = {}
        for lanes in candidates:
            rpt = self.rows_per_tile or choose_rows_per_tile(len(rows), int(rows.max()), r, lanes)
            height = None
            for budget in budgets:
                plan = native.PcgPlan(rows, r, lanes, rpt, ptr_off, idx_off, val_off, None, (indptr_h, indices_h),
                                      lds_bytes=budget)
                try:
                    if height is not None and (height >= rpt or plan.rows_per_tile <= height):
                        continue  # the wider budget stages no taller tile: same kernels
                    height = plan.rows_per_tile
                    times[(lanes, budget)] = self._time_iteration(plan, dev_ptr, dev_idx, values_d, device)
                finally:
                    plan.close()
        self.tuning[shape] = times
        return min(times, key=times.get)

    def _time_iteration(self, plan: native.PcgPlan, dev_ptr, dev_idx, values_d, device) -> float:
        """Seconds per iteration of the solver's preconditioner on ``plan`` (a tolerance below
        reach: exactly the requested number of iterations)."""

        precond = native.PRECOND_JACOBI
        if self.preconditioner in ("amg", "learned"):
            try:
                plan.attach_amg(values_d, theta=self.amg_theta, coarse_max=self.amg_coarse_max,
                                stall_limit=0, smoother_scale=self.amg_smoother_scale,
                                prolongator=self._fixed_prolongator(plan))
                precond = native.PRECOND_AMG
            except ValueError:  # no hierarchy for this pattern: time Jacobi PCG instead
                pass
        gen = torch.Generator(device=device).manual_seed(0)
        b = torch.randn(plan.vec_len, dtype=torch.float64, device=device, generator=gen)
        x = torch.zeros_like(b)

        def run(iterations):
            x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rc, _, _ = plan.solve(dev_ptr, dev_idx, values_d, b, x, 1e-300, iterations, precond=precond,
                                  check_every=8)
            torch.cuda.synchronize()
            if rc not in (native.ADT_OK, native.ADT_ERR_NOT_CONVERGED):
                native.check(rc)
            return time.perf_counter() - t0

        run(8)  # graphs and hierarchy values
        return (run(24) - run(8)) / 16

    def _fixed_prolongator(self, plan: native.PcgPlan):
        """``first_prolongator`` when it fits ``plan`` (one CSR pattern of its row count), else None."""

        fp = self.first_prolongator
        if fp is None or len(set(zip(plan.ptr_off.tolist(), plan.idx_off.tolist(), plan.rows.tolist()))) != 1:
            return None
        return fp if int(plan.rows[0]) == len(fp[0]) - 1 else None

    def _precond(self, plan: native.PcgPlan, values, preconditioner):
        """(native preconditioner id, name used, per-slot damping array or None)."""

        name = self.preconditioner if preconditioner is None else preconditioner
        if name == "jacobi":
            return native.PRECOND_JACOBI, "jacobi", None
        if name == "ic0":
            return None, "ic0", None
        if name not in ("amg", "learned"):
            raise ValueError(f"preconditioner must be one of {PRECONDITIONERS}, got {name!r}")
        fixed = self._fixed_prolongator(plan)
        if name == "learned" and fixed is not None:
            name = "amg"  # the network damps an aggregation prolongator; this plan's finest one is given
        if plan.amg_levels == 0 and not getattr(plan, "amg_failed", False):
            try:
                plan.attach_amg(values, theta=self.amg_theta, coarse_max=self.amg_coarse_max,
                                stall_limit=self.amg_stall_limit, smoother_scale=self.amg_smoother_scale,
                                prolongator=fixed)
            except ValueError:
                # no usable hierarchy (tiny or unstructured matrices): plain Jacobi PCG for this plan
                plan.amg_failed = True
        if plan.amg_levels == 0:
            return native.PRECOND_JACOBI, "jacobi", None
        if name == "amg":
            return native.PRECOND_AMG, "amg", None
        return native.PRECOND_LEARNED, "learned", self._learned_damping(plan, values)

    def _learned_damping(self, plan: native.PcgPlan, values):
        """Per-slot finest prolongation damping from the trained network, laid out like ``values``
        (the network runs once per distinct CSR pattern of the plan, on that pattern's systems)."""

        from . import learned

        if self.damping_network is None:
            if not learned.DEFAULT_WEIGHTS.exists():
                raise RuntimeError(f"no trained damping network at {learned.DEFAULT_WEIGHTS}")
            self.damping_network = learned.DampingNetwork.load()
        groups = getattr(plan, "gnn_groups", None)
        if groups is None:
            groups = self._damping_groups(plan, values.device, self.amg_theta)
            plan.gnn_groups = groups
        flat = values.reshape(-1)
        if len(groups) == 1 and groups[0]["contiguous"]:
            g = groups[0]
            v0 = int(plan.val_off[g["systems"][0]])
            omega = self.damping_network.omega(g["graph"], flat[v0:], g["omega_p"], val_stride=g["graph"].nnz)
            if v0 == 0:
                return omega
            out = torch.empty_like(flat)
            out[v0:v0 + omega.numel()] = omega.reshape(-1)
            return out
        out = torch.empty_like(flat)
        for g in groups:
            nnz = g["graph"].nnz
            offs = [int(plan.val_off[s]) for s in g["systems"]]
            stacked = torch.stack([flat[o:o + nnz] for o in offs]).reshape(-1)
            omega = self.damping_network.omega(g["graph"], stacked, g["omega_p"], val_stride=nnz).reshape(len(offs), nnz)
            for j, o in enumerate(offs):
                out[o:o + nnz] = omega[j]
        return out

    @staticmethod
    def _damping_groups(plan: native.PcgPlan, device, theta: float) -> list[dict]:
        """The plan's systems grouped by CSR pattern, each with its network graph (aggregation of
        its representative matrix) and the cycle's constant finest damping of every system."""

        from . import learned

        keyed: dict[tuple[int, int, int], list[int]] = {}
        for s in range(plan.n_sys):
            keyed.setdefault((int(plan.ptr_off[s]), int(plan.idx_off[s]), int(plan.rows[s])), []).append(s)
        rep_values, rep_off = plan.amg_representative
        groups = []
        for (p0, i0, n), systems in keyed.items():
            indptr = plan.host_pattern[0][p0:p0 + n + 1].astype(np.int64)
            nnz = int(indptr[-1] - indptr[0])
            indptr = (indptr - indptr[0]).astype(np.int32)
            indices = plan.host_pattern[1][i0:i0 + nnz]
            r0 = int(rep_off[systems[0]])
            agg = native.amg_aggregates(indptr, indices, rep_values[r0:r0 + nnz], theta=theta)
            graph = learned.GnnGraph.from_pattern(indptr, indices, agg, device)
            omega_p = torch.as_tensor([plan.amg_level(0, s)["omega"] for s in systems], dtype=torch.float64,
                                      device=device)
            offs = np.asarray([int(plan.val_off[s]) for s in systems])
            contiguous = bool(np.array_equal(offs, offs[0] + np.arange(len(systems)) * nnz))
            groups.append(dict(graph=graph, systems=systems, omega_p=omega_p, contiguous=contiguous))
        return groups

    def solve(self, plan: native.PcgPlan, indptr, indices, values, b, x, *, tol=None, max_iter=None,
              preconditioner=None, raise_on_failure=True):
        """Solve in place: ``x`` holds the initial guess on entry and the solution on return."""

        tol = self.tolerance if tol is None else tol
        max_iter = self.max_iterations if max_iter is None else max_iter
        precond, used, data = self._precond(plan, values, preconditioner)
        if used == "ic0":
            return self._solve_ic0(plan, indptr, indices, values, b, x, tol, max_iter, raise_on_failure)
        t0 = time.perf_counter()
        rc, iters, relres = plan.solve(indptr, indices, values, b, x, tol, max_iter, precond=precond,
                                       precond_data=data, check_every=self.check_every)
        st = plan.stats()
        fallback = st["fallback_rhs"] if precond != native.PRECOND_JACOBI else 0
        self.last = SolveStats(iters, relres, time.perf_counter() - t0, used, fallback, st["polished_rhs"])
        self.last_plan = plan
        if raise_on_failure and rc == native.ADT_ERR_NONFINITE:
            raise FloatingPointError("conjugate gradients produced non-finite values")
        if raise_on_failure and rc == native.ADT_ERR_NOT_CONVERGED:
            raise RuntimeError(f"conjugate gradients did not reach tol={tol} in {max_iter} iterations "
                               f"(worst relative residual {float(relres.max()):.3e})")
        if rc not in (native.ADT_OK, native.ADT_ERR_NONFINITE, native.ADT_ERR_NOT_CONVERGED):
            native.check(rc)
        return self.last

    def _solve_ic0(self, plan: native.PcgPlan, indptr, indices, values, b, x, tol, max_iter, raise_on_failure):
        """IC(0)-preconditioned CG (``adt_ic0_pcg``): the systems must share one CSR pattern, with
        values ``nnz`` apart and interleaved blocks one after another; a right-hand side shared by
        several systems (``b_off`` cycling through blocks) is kept shared."""

        n = int(plan.rows[0])
        nnz = int(plan.host_pattern[0][n]) if plan.host_pattern is not None else int(indices.numel())
        same = (np.all(plan.rows == n) and np.all(plan.ptr_off == 0) and np.all(plan.idx_off == 0)
                and np.array_equal(plan.val_off, np.arange(plan.n_sys) * nnz)
                and np.array_equal(plan.vec_off, np.arange(plan.n_sys) * n * plan.rp)
                and int(indptr.numel()) == n + 1 and int(indices.numel()) == nnz)
        if not same:
            raise ValueError("the IC(0) preconditioner needs systems that share one CSR pattern")
        block = n * plan.rp
        if plan.b_off is None:
            n_b = plan.n_sys
        else:
            n_b = int(plan.b_off.max() // block) + 1
            if not np.array_equal(plan.b_off, (np.arange(plan.n_sys) % n_b) * block) or plan.n_sys % n_b:
                raise ValueError("the IC(0) preconditioner needs right-hand sides cycling through whole blocks")
        patterns = getattr(plan, "ic0_device", None)
        if patterns is None:
            host_ptr = plan.host_pattern[0][: n + 1] if plan.host_pattern is not None else indptr[: n + 1].cpu().numpy()
            host_idx = plan.host_pattern[1][:nnz] if plan.host_pattern is not None else indices[:nnz].cpu().numpy()
            host = native.ic0_patterns(host_ptr, host_idx)
            patterns = {k: (torch.as_tensor(v, device=values.device) if isinstance(v, np.ndarray) else v)
                        for k, v in host.items()}
            plan.ic0_device = patterns
        t0 = time.perf_counter()
        rc, iters, relres = native.ic0_pcg(plan.n_sys, n, plan.r, plan.rp, indptr, indices, patterns,
                                           values.contiguous(), b.contiguous(), n_b, x, tol, max_iter,
                                           factor_sweeps=self.ic0_factor_sweeps, solve_sweeps=self.ic0_solve_sweeps,
                                           check_every=8)
        self.last = SolveStats(iters, relres, time.perf_counter() - t0, "ic0", 0, 0)
        self.last_plan = plan
        if raise_on_failure and rc == native.ADT_ERR_NONFINITE:
            raise FloatingPointError("conjugate gradients produced non-finite values")
        if raise_on_failure and rc == native.ADT_ERR_NOT_CONVERGED:
            raise RuntimeError(f"conjugate gradients did not reach tol={tol} in {max_iter} iterations "
                               f"(worst relative residual {float(relres.max()):.3e})")
        if rc not in (native.ADT_OK, native.ADT_ERR_NONFINITE, native.ADT_ERR_NOT_CONVERGED):
            native.check(rc)
        return self.last

    def close(self) -> None:
        for plan in self._plans.values():
            plan.close()
        self._plans.clear()


def _laplacian(indptr, indices, rows, ptr_off, idx_off):
    """Graph-Laplacian values (off-diagonals -1, diagonal = degree + 1e-3) for every system of a
    batch layout, one block per system: (values, val_off)."""

    blocks, val_off, offset = [], [], 0
    for s in range(len(rows)):
        p0, i0, n = int(ptr_off[s]), int(idx_off[s]), int(rows[s])
        ptr = indptr[p0:p0 + n + 1].astype(np.int64)
        nnz = int(ptr[-1])
        cols = indices[i0:i0 + nnz]
        row_of = np.repeat(np.arange(n), np.diff(ptr))
        vals = np.where(cols == row_of, 0.0, -1.0)
        deg = np.bincount(row_of, weights=-vals, minlength=n)
        vals = np.where(cols == row_of, deg[row_of] + 1e-3, vals)
        blocks.append(vals)
        val_off.append(offset)
        offset += nnz
    return np.concatenate(blocks), np.asarray(val_off, dtype=np.int64)


def interleave(fields: torch.Tensor, rp: int) -> torch.Tensor:
    """``(S, R, N)`` field rows -> ``(S, N, Rp)`` interleaved blocks (zero padding columns)."""

    s, r, n = fields.shape
    out = torch.zeros((s, n, rp), dtype=torch.float64, device=fields.device)
    out[:, :, :r] = fields.transpose(1, 2)
    return out
