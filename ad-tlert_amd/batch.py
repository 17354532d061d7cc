"""Several 2.5D operators on meshes of different sizes solved in one batched launch.

The reference batches the ``steps x wavenumbers`` systems of one mesh per direct solve
(``forward/ert2p5d.py:532-553``, ``forward/cudss.py:63-122``); meshes of different sizes are
separate operators and separate solves.  ``MixedMeshBatch`` gives a set of ``ERTForward2p5D``
operators one native plan: every (operator, time step, wavenumber) system, each with its own CSR
pattern, right-hand side and multigrid hierarchy, iterates in the same captured iteration graph,
so the small meshes fill the device around the large ones instead of running alone.  Sources are
padded to the largest operator's count (padding columns are zero and never iterate).

The results are each operator's own ``apparent_resistivity_values`` / ``resistance`` (same
solver tolerance, same measurement code), computed together.
"""

from __future__ import annotations

import numpy as np
import torch

from .dtypes import FLOAT_DTYPE
from .ert2p5d import ERTForward2p5D
from .inputs import check_finite
from .solver import BatchedSolver


class MixedMeshBatch:
    """One solver plan for the systems of several 2.5D operators (any mesh sizes)."""

    def __init__(self, operators, *, preconditioner: str | None = None, solver_tolerance: float | None = None):
        self.operators = list(operators)
        if not self.operators:
            raise ValueError("MixedMeshBatch needs at least one operator")
        if not all(isinstance(op, ERTForward2p5D) for op in self.operators):
            raise TypeError("MixedMeshBatch batches ERTForward2p5D operators")
        first = self.operators[0]._solver
        self._solver = BatchedSolver(tolerance=first.tolerance if solver_tolerance is None else float(solver_tolerance),
                                     preconditioner=first.preconditioner if preconditioner is None else preconditioner)
        self._layout = None

    @property
    def solver(self) -> BatchedSolver:
        return self._solver

    def _build_layout(self, steps: int):
        """Concatenated CSR patterns (host and device) and per-system offsets for ``steps``
        models per operator: system order (operator, step, wavenumber)."""

        devs = [op._dev for op in self.operators]
        r = max(d.E for d in devs)
        ptr_parts, idx_parts, ptr_base, idx_base = [], [], [], []
        p = i = 0
        for d in devs:
            ptr_base.append(p)
            idx_base.append(i)
            ptr_parts.append(d.pattern.indptr.astype(np.int32))
            idx_parts.append(d.pattern.indices.astype(np.int32))
            p += d.N + 1
            i += d.nnz
        rows, ptr_off, idx_off, val_off, owner = [], [], [], [], []
        v = 0
        for k, d in enumerate(devs):
            for _ in range(steps):
                for _ in range(d.W):
                    rows.append(d.N)
                    ptr_off.append(ptr_base[k])
                    idx_off.append(idx_base[k])
                    val_off.append(v)
                    owner.append(k)
                    v += d.nnz
        host_pattern = (np.concatenate(ptr_parts), np.concatenate(idx_parts))
        device = devs[0].device
        self._layout = dict(
            steps=steps, r=r, rows=np.asarray(rows, np.int32), offsets=(np.asarray(ptr_off, np.int64),
                                                                        np.asarray(idx_off, np.int64),
                                                                        np.asarray(val_off, np.int64)),
            owner=np.asarray(owner), host_pattern=host_pattern, values_len=v,
            indptr=torch.as_tensor(host_pattern[0], device=device), indices=torch.as_tensor(host_pattern[1], device=device))
        return self._layout

    def _plan(self, steps: int):
        layout = self._layout if self._layout is not None and self._layout["steps"] == steps else self._build_layout(steps)
        plan = self._solver.plan(("mixed", steps), layout["rows"], layout["r"], pattern_offsets=layout["offsets"],
                                 host_pattern=layout["host_pattern"], lanes=None)
        return layout, plan

    def solve_fields(self, conductivities):
        """Interleaved total fields ``(steps, W, N, Rp)`` of every operator for its models
        (``conductivities[k]`` is one model ``(C_k,)`` or a series ``(steps, C_k)``; every operator
        gets the same number of steps)."""

        models = []
        for op, sigma in zip(self.operators, conductivities, strict=True):
            s = torch.as_tensor(sigma)
            models.append(op._series(s if s.ndim == 2 else s[None]))
        steps = int(models[0].shape[0])
        if any(int(m.shape[0]) != steps for m in models):
            raise ValueError("every operator needs the same number of models")
        layout, plan = self._plan(steps)
        device = self.operators[0]._dev.device
        values = torch.empty(layout["values_len"], dtype=torch.float64, device=device)
        x = torch.zeros(plan.vec_len, dtype=torch.float64, device=device)
        b = torch.zeros(plan.vec_len, dtype=torch.float64, device=device)
        s = 0
        rp = plan.rp
        for op, sig in zip(self.operators, models):
            d = op._dev
            sig = sig.to(device=device, dtype=torch.float64)
            vals = op._assemble_device(d, sig)  # (steps * W, nnz)
            rho = op._source_resistivities(d, sig)  # (steps, E)
            for t in range(steps):
                for w in range(d.W):
                    o, n = int(plan.vec_off[s]), d.N
                    v0 = int(layout["offsets"][2][s])
                    values[v0:v0 + d.nnz] = vals[t * d.W + w]
                    block_x = x[o:o + n * rp].view(n, rp)
                    block_b = b[o:o + n * rp].view(n, rp)
                    block_x[:, : d.E] = d.primary[w, :, : d.E] * rho[t][None, :]
                    block_b[:, : d.E] = d.rhs[w, :, : d.E]
                    s += 1
        self._solver.solve(plan, layout["indptr"], layout["indices"], values, b, x)
        fields, s = [], 0
        for op in self.operators:
            d = op._dev
            blocks = torch.empty((steps, d.W, d.N, rp), dtype=torch.float64, device=device)
            for t in range(steps):
                for w in range(d.W):
                    o = int(plan.vec_off[s])
                    blocks[t, w] = x[o:o + d.N * rp].view(d.N, rp)
                    s += 1
            fields.append(blocks)
        return fields

    def resistance(self, conductivities) -> list[torch.Tensor]:
        """Each operator's ``resistance`` (one model) or ``resistance_series`` (a series)."""

        fields = self.solve_fields(conductivities)
        out = []
        for op, blocks, sigma in zip(self.operators, fields, conductivities):
            potentials = op._potentials(blocks)
            check_finite(potentials, "integrated potentials")
            r = op._combine(*op._normal_reciprocal(potentials)).cpu().to(FLOAT_DTYPE)
            out.append(r if torch.as_tensor(sigma).ndim == 2 else r[0])
        return out

    def apparent_resistivity_values(self, conductivities, currents=1.0) -> list[torch.Tensor]:
        current = torch.as_tensor(currents, dtype=FLOAT_DTYPE)
        return [op._geometric_factors() * r / current for op, r in zip(self.operators, self.resistance(conductivities))]

    @property
    def rhs_per_solve(self) -> int:
        """Right-hand sides of one model per operator (wavenumbers x sources, summed)."""

        return sum(op._dev.W * op._dev.E for op in self.operators)

    def close(self) -> None:
        self._solver.close()

