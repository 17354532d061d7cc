"""Three-dimensional DC ERT forward operator on tetrahedral P1/P2 meshes, on one MI355X.

Drop-in for ``adtlert.forward.ERTForward3D`` (``forward/ert3d.py:114-983``): pyGIMLi's
singularity-removal (SR) formulation on flat surfaces (``A(sigma) u = A(1) u_p`` with the
analytic half-space primary ``u_p``), point sources otherwise; mixed (far-field Robin) or
Dirichlet outer boundaries; analytic or numerical geometric factors.  Assembly, FP64 PCG
solves, electrode potentials, the fused sensitivity kernel and the adjoint gradient run in the
native gfx950 library.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import fem, native
from .ordering import dof_order, inverse
from .dtypes import FLOAT_DTYPE
from .ert2p5d import _Columns, chunked_rows, electrode_stencil, jacobian_rows_per_launch, slot_map
from .field_cache import FieldCache, digest16, newest_batches_first, step_batches
from .inputs import (cell_vector, check_finite, conductivity_model, conductivity_series, cotangent_series,
                     measurement_vector, option, parameter_targets)
from .mesh import TET_EDGES, Mesh3D
from .solver import BatchedSolver, choose_lanes
from .survey import Survey

Tensor = torch.Tensor
PYGIMLI_ELECTRODE_RADIUS_FRACTION = 0.5
"""pyGIMLi's primary potential at an electrode node uses this fraction of the shortest distance
to another node (``ert3d.py:47``)."""


@dataclass
class _Fields3D:
    """Assembled values, interleaved current-electrode fields and (once solved) receiver fields."""

    values: Tensor
    current: Tensor
    receiver: Tensor | None = None


@dataclass(frozen=True)
class ForwardResponse3D:
    apparent_resistivity: Tensor
    resistance: Tensor
    electrode_potentials: Tensor | None = None
    node_potentials: Tensor | None = None


def _electrode_rows(slots: np.ndarray, count: int):
    """Electrodes used in ``slots`` and the survey-slot -> field-row map (remote -> one past the last)."""

    ids = np.unique(slots)
    ids = ids[ids < count].astype(np.int64)
    rows = np.full(count + 1, -1, dtype=np.int64)
    rows[ids] = np.arange(ids.size)
    rows[count] = ids.size
    return ids, rows


def pair_scatter_3d(a, b, m, n, n_curr, n_recv, cotangents: np.ndarray) -> np.ndarray:
    """``(Q, D)`` cotangents -> ``(Q, n_curr, n_recv)`` (``ert3d.py:688-706``), sequential host scatter."""

    width = n_recv + 1
    g = np.atleast_2d(np.asarray(cotangents, dtype=np.float64))
    q = g.shape[0]
    pairs = np.zeros((q, (n_curr + 1) * width))
    rows = np.broadcast_to(np.arange(q)[:, None], g.shape)
    for idx, sign in ((a * width + m, 1.0), (a * width + n, -1.0), (b * width + m, -1.0), (b * width + n, 1.0)):
        np.add.at(pairs, (rows, np.broadcast_to(idx, g.shape)), sign * g)
    return pairs.reshape(q, n_curr + 1, width)[:, :n_curr, :n_recv]


class ERTForward3D:
    """Tetrahedral P1/P2 forward operator (``ert3d.py:114-218``)."""

    def __init__(self, mesh: Mesh3D, survey: Survey, boundary_mode: str = "mixed", singularity_removal: bool = True,
                 element_order: int = 1, geometric_factor_mode: str = "auto", normal_field_cache_max_entries: int = 8,
                 series_batch_steps: int = 1, *, solver_tolerance: float = 1e-10, sensitivity_precision: str = "float64",
                 preconditioner: str = "amg"):
        self.mesh, self.survey = mesh, survey
        self.boundary_mode = option(boundary_mode, ("mixed", "dirichlet"), "boundary_mode")
        self.geometric_factor_mode = option(geometric_factor_mode, ("auto", "analytic", "numerical"),
                                            "geometric_factor_mode")
        if not torch.cuda.is_available():
            raise RuntimeError("ADTLERT-MI355X requires an AMD GPU with HIP (no device is visible)")
        if element_order not in (1, 2):
            raise ValueError("element_order must be 1 or 2")
        if survey.dimension != 3:
            raise ValueError("ERTForward3D requires three-coordinate electrode positions")
        if int(series_batch_steps) < 1:
            raise ValueError("series_batch_steps must be >= 1")
        self.singularity_removal = bool(singularity_removal)
        self.element_order = int(element_order)
        self.series_batch_steps = int(series_batch_steps)
        self._field_cache = FieldCache(normal_field_cache_max_entries)
        self.normal_field_cache_max_entries = self._field_cache.max_entries
        self.sensitivity_precision = option(sensitivity_precision, ("float64", "mixed"), "sensitivity_precision")
        self._solver = BatchedSolver(tolerance=float(solver_tolerance),
                                     preconditioner=option(preconditioner, ("amg", "jacobi", "learned", "ic0"), "preconditioner"))
        self._cache: dict = {}
        electrodes = np.asarray(survey.electrode_positions, dtype=float)
        quads = survey.measurement_slots().numpy()
        self._slots = quads
        self._current_electrode_ids, self._current_electrode_map = _electrode_rows(quads[:, :2], survey.electrode_count)
        self._receiver_electrode_ids, self._receiver_electrode_map = _electrode_rows(quads[:, 2:], survey.electrode_count)
        nodes = np.asarray(mesh.nodes, dtype=float)
        cells = np.asarray(mesh.cells, dtype=np.int64)
        if self.element_order == 2:
            dof_nodes, conn, bconn = mesh.build_quadratic_topology()
            self._dof_nodes = np.asarray(dof_nodes, dtype=float)
            self._cell_connectivity = np.asarray(conn, dtype=np.int64)
            self._boundary_connectivity = np.asarray(bconn, dtype=np.int64)
            self._volume_templates = fem.tet_p2(nodes, cells)
        else:
            self._dof_nodes = nodes
            self._cell_connectivity = cells
            self._boundary_connectivity = np.asarray(mesh.boundary_faces, dtype=np.int64)
            self._volume_templates = fem.tet_p1(nodes, cells)
        self.N = self._dof_nodes.shape[0]
        cell_ids, bary = mesh.locate_points(electrodes)
        weights = np.asarray(bary, dtype=float)
        if self.element_order == 2:
            weights = fem.tet_p2_shape(weights)
        k = self._cell_connectivity.shape[1]
        self._electrode_nodes = self._cell_connectivity[np.asarray(cell_ids)]  # (E, K)
        self._electrode_weights = weights  # (E, K)
        self._cell_templates = self._volume_templates.copy()
        if self.boundary_mode == "mixed":
            fem.add_robin_faces(self._cell_templates, self._cell_connectivity, self._boundary_connectivity,
                                np.asarray(mesh.boundary_face_cells), np.asarray(mesh.boundary_face_centers),
                                np.asarray(mesh.boundary_face_areas), np.asarray(mesh.boundary_face_normals),
                                np.asarray(mesh.surface_face_mask), electrodes.mean(axis=0), self.element_order)
        self._shared_source_receiver_rhs = bool((not self.singularity_removal or not mesh.is_flat_surface)
                                                and np.array_equal(self._current_electrode_ids,
                                                                   self._receiver_electrode_ids))
        self._dev_state = None
        del k

    @classmethod
    def from_mesh_survey(cls, mesh: Mesh3D, survey: Survey, **options) -> "ERTForward3D":
        return cls(mesh=mesh, survey=survey, **options)

    @property
    def cell_count(self) -> int:
        return self.mesh.cell_count

    # -- host precomputation ---------------------------------------------------------------------

    def _dirichlet_node_ids(self) -> np.ndarray:
        far = self._boundary_connectivity[~np.asarray(self.mesh.surface_face_mask, dtype=bool)]
        return np.unique(far.reshape(-1))

    def _assembly_routing(self, conn: np.ndarray, fixed: np.ndarray):
        """Local-entry -> CSR slot routing for connectivity ``conn`` with the Dirichlet nodes
        ``fixed`` masked (``ert3d.py:290-351``): CSR arrays, per-entry slot (-1 masked), unit
        diagonal slots and the masked flat templates."""

        k = conn.shape[1]
        n = self.N
        C = conn.shape[0]
        rows = np.repeat(conn, k, axis=1).reshape(-1)
        cols = np.tile(conn, (1, k)).reshape(-1)
        keys = rows * n + cols
        if fixed.size:
            mask = np.zeros(n, dtype=bool)
            mask[fixed] = True
            volume_pair = ~mask[rows] & ~mask[cols]
            unique = np.unique(keys[volume_pair | (rows == cols)])
            slots = np.searchsorted(unique, keys)
            groups = np.where(volume_pair, slots, -1)
            fixed_slots = np.searchsorted(unique, fixed * n + fixed)
            pm = volume_pair.reshape(C, -1)
            flat = np.where(pm, self._cell_templates.reshape(C, -1), 0.0)
            flat_v = np.where(pm, self._volume_templates.reshape(C, -1), 0.0)
        else:
            unique, inv = np.unique(keys, return_inverse=True)
            groups = inv.reshape(-1)
            fixed_slots = np.zeros(0, dtype=np.int64)
            flat = self._cell_templates.reshape(C, -1)
            flat_v = self._volume_templates.reshape(C, -1)
        indices = (unique % n).astype(np.int32)
        indptr = np.concatenate(([0], np.cumsum(np.bincount(unique // n, minlength=n)))).astype(np.int32)
        return indptr, indices, groups, fixed_slots, np.ascontiguousarray(flat), np.ascontiguousarray(flat_v)

    def _primary(self, positions: np.ndarray) -> np.ndarray:
        """``(n, N)`` half-space primaries ``(1/r + 1/r')/(4 pi)`` with the node fallback radius (``ert3d.py:367-385``)."""

        nodes = self._dof_nodes
        level = float(self.mesh.surface_reference_level)
        out = np.empty((positions.shape[0], nodes.shape[0]))
        for i, src in enumerate(positions):
            mirror = src.copy()
            mirror[2] = 2.0 * level - src[2]
            dd = np.linalg.norm(nodes - src, axis=1)
            md = np.linalg.norm(nodes - mirror, axis=1)
            fallback = max(float(np.min(dd[dd > 1e-12])) * PYGIMLI_ELECTRODE_RADIUS_FRACTION, 1e-12)
            dd = np.where(dd > 1e-12, dd, fallback)
            md = np.where(md > 1e-12, md, fallback)
            out[i] = (1.0 / dd + 1.0 / md) / (4.0 * np.pi)
        return out

    # -- device state ----------------------------------------------------------------------------

    @property
    def _dev(self):
        if self._dev_state is None:
            self._dev_state = self._build_device()
        return self._dev_state

    def _build_device(self):
        """Device state in a Hilbert-curve DOF numbering (``perm``: device row -> mesh DOF)."""

        dev = native.device()
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        s = type("Dev3D", (), {})()
        s.device = dev
        perm = dof_order(self._dof_nodes)
        inv = inverse(perm)
        s.perm_host, s.inv_host = perm, inv
        s.inv = torch.as_tensor(inv, dtype=torch.long, device=dev)
        conn = inv[self._cell_connectivity]
        fixed_mesh = self._dirichlet_node_ids() if self.boundary_mode == "dirichlet" else np.zeros(0, np.int64)
        fixed_nodes = np.sort(inv[fixed_mesh])
        indptr, indices, groups, fixed_slots, flat, flat_v = self._assembly_routing(conn, fixed_nodes)
        self.nnz = int(indices.size)
        s.indptr = torch.as_tensor(indptr, **i32)
        s.indices = torch.as_tensor(indices, **i32)
        s.host_pattern = (indptr, indices)
        sptr, sent = slot_map(groups, self.nnz)
        s.fixed = torch.as_tensor(fixed_slots.astype(np.int32), **i32) if fixed_slots.size else None
        k = conn.shape[1]
        C = conn.shape[0]
        s.templates = torch.as_tensor(flat.reshape(C, k, k), **f64).contiguous()
        s.volume_templates = torch.as_tensor(flat_v.reshape(C, k, k), **f64).contiguous()
        s.cell_col = torch.arange(C, **i32)
        # slot-ordered scatter map with both template sets laid out in its order
        s.amap = native.AssemblyMap(sptr, sent, C * k * k, k * k, np.arange(C), device=dev)
        s.etemplates = s.amap.entries(s.templates)
        s.evolume_templates = s.amap.entries(s.volume_templates)
        s.cell_dofs = torch.as_tensor(conn, **i32).contiguous()
        s.cell_dofs_mesh = torch.as_tensor(self._cell_connectivity, **i32).contiguous()
        s.N, s.nnz = self.N, self.nnz
        nc, nr = self._current_electrode_ids.size, self._receiver_electrode_ids.size
        s.nc, s.nr = nc, nr
        s.lanes_c, s.lanes_r = choose_lanes(nc), choose_lanes(nr)
        s.rp_c, s.rp_r = -(-nc // s.lanes_c) * s.lanes_c, -(-nr // s.lanes_r) * s.lanes_r
        # receiver rows (Dirichlet-zeroed electrode interpolation rows) as stencils and dense
        en, ew = inv[self._electrode_nodes], self._electrode_weights.copy()
        if fixed_nodes.size:
            ew[np.isin(en, fixed_nodes)] = 0.0
        recv = self._receiver_electrode_ids
        s.enode = torch.as_tensor(en[recv].astype(np.int32), **i32)
        s.eweight = torch.as_tensor(ew[recv], **f64)
        dense = np.zeros((nr, self.N))
        np.add.at(dense, (np.repeat(np.arange(nr), en.shape[1]), en[recv].reshape(-1)), ew[recv].reshape(-1))
        s.receiver_rows_t = torch.as_tensor(dense.T.copy(), **f64)  # (N, n_recv)
        # right-hand sides (device numbering)
        positions = np.asarray(self.survey.electrode_positions, dtype=float)
        unit = self._assemble(s, torch.ones((1, C), **f64), robin=True, fixed=True)
        s.unit_values = unit
        prim_c = self._primary(positions[self._current_electrode_ids])[:, perm]
        prim_r = self._primary(positions[recv])[:, perm]
        s.primary_c = self._interleave(prim_c, s.rp_c, dev)
        s.primary_r = self._interleave(prim_r, s.rp_r, dev)
        if self.singularity_removal and self.mesh.is_flat_surface:
            s.rhs_c = torch.empty_like(s.primary_c)
            self._spmm(s, unit, s.primary_c, s.rhs_c)
        else:
            point = np.zeros((nc, self.N))
            cur = self._current_electrode_ids
            np.add.at(point, (np.repeat(np.arange(nc), en.shape[1]), en[cur].reshape(-1)), ew[cur].reshape(-1))
            s.rhs_c = self._interleave(point, s.rp_c, dev)
        if fixed_nodes.size:
            s.rhs_c[torch.as_tensor(fixed_nodes, device=dev)] = 0.0
        s.rhs_r = self._interleave(dense, s.rp_r, dev)
        s.fixed_nodes = torch.as_tensor(fixed_nodes, device=dev) if fixed_nodes.size else None
        # warm-start resistivity around every electrode node
        cells = self._cell_connectivity[:, :4]
        nearest = np.argmin(np.linalg.norm(self._dof_nodes[None] - positions[:, None], axis=2), axis=1)
        lists = [np.flatnonzero(np.any(cells == node, axis=1)) for node in nearest]
        width = max(1, max(len(x) for x in lists))
        table = np.full((positions.shape[0], width), C, dtype=np.int64)
        for e, x in enumerate(lists):
            table[e, : len(x)] = x
        s.src_table = torch.as_tensor(table, device=dev)
        s.src_count = torch.as_tensor(np.maximum([len(x) for x in lists], 1), **f64)
        a, b = self._current_electrode_map[self._slots[:, 0]], self._current_electrode_map[self._slots[:, 1]]
        m, n = self._receiver_electrode_map[self._slots[:, 2]], self._receiver_electrode_map[self._slots[:, 3]]
        s.abmn_host = (a, b, m, n)
        s.abmn = [torch.as_tensor(x, dtype=torch.long, device=dev) for x in (a, b, m, n)]
        s.abmn_i32 = [torch.as_tensor(x.astype(np.int32), device=dev) for x in (a, b, m, n)]
        if self.element_order == 2:
            self._solver.first_prolongator = self._vertex_prolongator(inv)
        return s

    def _vertex_prolongator(self, inv: np.ndarray) -> tuple:
        """Interpolation of the ten-node DOFs from the vertices in the device numbering (rows: DOFs,
        columns: vertices ordered by their device row): a vertex copies itself, an edge midpoint
        takes half of each end.  The multigrid cycle's finest prolongator on P2 operators: point
        smoothing reduces the P2-specific error components poorly, and this interpolation hands
        them to a P1-like Galerkin operator (14 entries per row on the terrain mesh instead of ~70
        on the first aggregation level), which aggregates on."""

        nv = int(self.mesh.node_count)
        conn = self._cell_connectivity
        mid = conn[:, 4:].reshape(-1)
        ends = np.asarray(TET_EDGES)
        a = conn[:, ends[:, 0]].reshape(-1)
        b = conn[:, ends[:, 1]].reshape(-1)
        mid, first = np.unique(mid, return_index=True)
        col = np.empty(nv, dtype=np.int64)
        col[np.argsort(inv[:nv], kind="stable")] = np.arange(nv)
        rows = np.concatenate((inv[:nv], inv[mid], inv[mid]))
        cols = np.concatenate((col, col[a[first]], col[b[first]]))
        vals = np.concatenate((np.ones(nv), np.full(2 * mid.size, 0.5)))
        order = np.lexsort((cols, rows))
        indptr = np.concatenate(([0], np.cumsum(np.bincount(rows, minlength=self.N)))).astype(np.int32)
        return indptr, cols[order].astype(np.int32), vals[order], nv

    @staticmethod
    def _interleave(rows: np.ndarray, rp: int, dev) -> Tensor:
        out = torch.zeros((rows.shape[1], rp), dtype=torch.float64, device=dev)
        out[:, : rows.shape[0]] = torch.as_tensor(rows.T, dtype=torch.float64, device=dev)
        return out

    def _assemble(self, s, sigmas: Tensor, *, robin: bool, fixed: bool) -> Tensor:
        """``(T, nnz)`` values of ``sum_c sigma_c T_c`` (+ unit Dirichlet diagonals) (``ert3d.py:436-461``)."""

        sigmas = sigmas.to(device=s.device, dtype=torch.float64).contiguous()
        values = torch.empty((sigmas.shape[0], self.nnz), dtype=torch.float64, device=s.device)
        native.assemble(s.amap, sigmas, s.etemplates if robin else s.evolume_templates, values, n_waves=1,
                        fixed=s.fixed if fixed else None)
        return values

    def _spmm(self, s, values: Tensor, x: Tensor, y: Tensor, mode: int = 0) -> None:
        n_sys = values.shape[0]
        rp = x.shape[-1]
        rows = np.full(n_sys, self.N, dtype=np.int32)
        z = np.zeros(n_sys, dtype=np.int64)
        layout = (rows, z, z, np.arange(n_sys, dtype=np.int64) * self.nnz, np.arange(n_sys, dtype=np.int64) * self.N * rp,
                  1, rp)
        native.spmm(layout, s.indptr, s.indices, values.contiguous(), x.contiguous(), y, mode)

    def _pcg(self, s, key, values: Tensor, rhs: Tensor, x: Tensor, r: int, shared_rhs: bool = False) -> None:
        n_sys = int(values.shape[0])
        rows = np.full(n_sys, self.N, dtype=np.int32)
        val_off = np.arange(n_sys, dtype=np.int64) * self.nnz
        b_off = np.zeros(n_sys, dtype=np.int64) if shared_rhs else None
        plan = self._solver.plan((key, n_sys), rows, r, pattern_offsets=(np.zeros(n_sys, np.int64),
                                 np.zeros(n_sys, np.int64), val_off), b_off=b_off, host_pattern=s.host_pattern)
        if plan.rp != x.shape[-1] or not np.array_equal(plan.vec_off, np.arange(n_sys) * self.N * plan.rp):
            raise RuntimeError("unexpected PCG vector layout")
        self._solver.solve(plan, s.indptr, s.indices, values.contiguous(), rhs.contiguous(), x)

    def _warm_start(self, s, sigmas: Tensor, primary: Tensor, ids: np.ndarray) -> Tensor:
        logs = torch.nn.functional.pad(-torch.log(sigmas), (0, 1))
        rho = torch.exp(logs[:, s.src_table].sum(dim=-1) / s.src_count)[:, torch.as_tensor(ids, device=s.device)]
        x = primary[None] * torch.nn.functional.pad(rho, (0, primary.shape[-1] - rho.shape[1]))[:, None, :]
        if s.fixed_nodes is not None:
            x[:, s.fixed_nodes] = 0.0
        return x.contiguous()

    # -- fields ----------------------------------------------------------------------------------

    def _sigma(self, conductivity) -> Tensor:
        return conductivity_model(conductivity, self.mesh.cell_count, torch.float64)

    def _series(self, conductivities) -> Tensor:
        return conductivity_series(conductivities, self.mesh.cell_count, torch.float64)

    def _solve_fields(self, sigmas: Tensor) -> list[_Fields3D]:
        s = self._dev
        sigmas = sigmas.to(device=s.device, dtype=torch.float64)
        values = self._assemble(s, sigmas, robin=True, fixed=True)
        x = self._warm_start(s, sigmas, s.primary_c, self._current_electrode_ids)
        self._pcg(s, "current", values, s.rhs_c, x, s.nc, shared_rhs=True)
        out = []
        for i in range(sigmas.shape[0]):
            f = _Fields3D(values[i:i + 1], x[i])
            if self._shared_source_receiver_rhs:
                f.receiver = f.current
            out.append(f)
        return out

    def _fields_series(self, models: Tensor):
        return self._field_cache.solve_series(models, self.series_batch_steps, self._solve_fields)

    def _fields(self, conductivity) -> _Fields3D:
        return self._fields_series(self._sigma(conductivity)[None])[0][0]

    def _receivers(self, fields: list[_Fields3D], models: Tensor) -> list[Tensor]:
        s = self._dev
        missing = [i for i, f in enumerate(fields) if f.receiver is None]
        for batch in step_batches(missing, self.series_batch_steps)[::-1]:
            values = torch.cat([fields[i].values for i in batch])
            sig = models[batch].to(device=s.device, dtype=torch.float64)
            x = self._warm_start(s, sig, s.primary_r, self._receiver_electrode_ids)
            self._pcg(s, "receiver", values, s.rhs_r, x, s.nr, shared_rhs=True)
            for j, i in enumerate(batch):
                fields[i].receiver = x[j]
        return [f.receiver for f in fields]

    def normal_field_cache_info(self) -> dict[str, int]:
        return self._field_cache.info()

    # -- measurement -----------------------------------------------------------------------------

    def _potentials(self, current: Tensor) -> Tensor:
        """``(T, n_curr, n_recv)`` potentials of interleaved current fields ``(T, N, Rp)``."""

        s = self._dev
        current = current.contiguous()
        T = current.shape[0]
        out = torch.empty((T, s.nc, s.nr), dtype=torch.float64, device=s.device)
        offs = np.arange(T, dtype=np.int64) * self.N * current.shape[-1]
        native.electrode_potentials(current, offs, T, 1, current.shape[-1], s.nc, s.enode, s.eweight, np.ones(1), out)
        return out

    def _measurement_resistances(self, current: Tensor) -> Tensor:
        s = self._dev
        p = torch.nn.functional.pad(self._potentials(current), (0, 1, 0, 1))
        a, b, m, n = s.abmn
        return p[..., a, m] - p[..., b, m] - p[..., a, n] + p[..., b, n]

    def _geometric_factors(self) -> Tensor:
        numerical = self.geometric_factor_mode == "numerical" or (
            self.geometric_factor_mode == "auto" and not self.mesh.is_flat_surface)
        if not numerical:
            return self.survey.geometric_factors()
        if "geometric_factors" not in self._cache:
            s = self._dev
            x = torch.zeros((1, self.N, s.rp_c), dtype=torch.float64, device=s.device)
            x[0] = s.primary_c
            self._pcg(s, "unit", s.unit_values, s.rhs_c, x, s.nc, shared_rhs=True)
            unit = self._measurement_resistances(x)[0]
            if bool(torch.any(torch.abs(unit) <= torch.finfo(torch.float64).tiny)):
                raise ValueError("numerical geometric-factor solve produced zero resistance")
            self._cache["geometric_factors"] = (1.0 / unit).to(device="cpu", dtype=FLOAT_DTYPE)
        return self._cache["geometric_factors"]

    def _per_electrode(self, rows: np.ndarray) -> Tensor:
        full = np.full((self.survey.electrode_count, rows.shape[1]), np.nan)
        full[self._current_electrode_ids] = rows
        return torch.as_tensor(full, dtype=FLOAT_DTYPE)

    def _response(self, fields: _Fields3D, currents, *, include_fields: bool) -> ForwardResponse3D:
        resistance = self._measurement_resistances(fields.current[None])[0].to(device="cpu", dtype=FLOAT_DTYPE)
        apparent = self._geometric_factors() * resistance / torch.as_tensor(currents, dtype=FLOAT_DTYPE)
        if not include_fields:
            return ForwardResponse3D(apparent, resistance)
        s = self._dev
        nodes = fields.current[s.inv, : s.nc].T.cpu().numpy()
        en, ew = self._electrode_nodes, self._electrode_weights
        at_electrodes = np.einsum("snk,ek->se", nodes[:, en], ew) if nodes.size else np.zeros((0, 0))
        return ForwardResponse3D(apparent, resistance, electrode_potentials=self._per_electrode(at_electrodes),
                                 node_potentials=self._per_electrode(nodes))

    def solve(self, conductivity, currents=1.0) -> ForwardResponse3D:
        return self._response(self._fields(conductivity), currents, include_fields=True)

    def resistance(self, conductivity) -> Tensor:
        return self.resistance_series(self._sigma(conductivity)[None])[0]

    def resistance_series(self, conductivities) -> Tensor:
        fields, _ = self._fields_series(self._series(conductivities))
        out = self._measurement_resistances(torch.stack([f.current for f in fields])).to(device="cpu", dtype=FLOAT_DTYPE)
        check_finite(out, "resistances")
        return out

    def apparent_resistivity_values(self, conductivity, currents=1.0) -> Tensor:
        return self._geometric_factors() * self.resistance(conductivity) / torch.as_tensor(currents, dtype=FLOAT_DTYPE)

    def apparent_resistivity_series(self, conductivities, currents=1.0) -> Tensor:
        factors = self._geometric_factors()
        resistance = self.resistance_series(conductivities)
        current = torch.as_tensor(currents, dtype=FLOAT_DTYPE)
        if current.ndim == 1 and current.shape[0] == resistance.shape[0] and current.shape[0] != resistance.shape[1]:
            current = current[:, None]
        return factors * resistance / current

    # -- derivatives -----------------------------------------------------------------------------

    def _columns(self, cell_parameter_ids=None, parameter_count=None) -> _Columns:
        targets, count = parameter_targets(cell_parameter_ids, parameter_count, self.mesh.cell_count)
        # the default (every cell its own column) needs no digest of its cell_count ids
        key = ("columns", None if cell_parameter_ids is None else digest16(targets.tobytes()), count)
        if key not in self._cache:
            self._cache[key] = _Columns(targets, count, self._dev.device)
        return self._cache[key]

    def _cell_gradient(self, phi: Tensor, phi_rp: int, lam: Tensor, lam_rp: int, r: int, *, robin: bool) -> Tensor:
        """``-sum_s lam_s^T T_c phi_s`` per cell (``ert3d.py:713-734``), FP64 on the device."""

        s = self._dev
        out = torch.empty(self.mesh.cell_count, dtype=torch.float64, device=s.device)
        native.cell_gradient(lam.contiguous(), lam_rp, [0], phi.contiguous(), phi_rp, [0], r, s.cell_dofs,
                             s.templates if robin else s.volume_templates, None, [1.0], None, out)
        return out

    def _pairs(self, cotangents: np.ndarray) -> np.ndarray:
        s = self._dev
        a, b, m, n = s.abmn_host
        return pair_scatter_3d(a, b, m, n, s.nc, s.nr, cotangents)

    def _combine_receivers(self, pairs: np.ndarray, receiver: Tensor, rp: int) -> Tensor:
        """``(N, rp)`` interleaved rows ``sum_f pairs[s, f] v_f`` of receiver fields or rows."""

        s = self._dev
        p = torch.as_tensor(pairs, dtype=torch.float64, device=s.device)  # (nc, nr)
        out = torch.zeros((self.N, rp), dtype=torch.float64, device=s.device)
        out[:, : s.nc] = receiver[:, : s.nr] @ p.T
        return out

    def vjp(self, conductivity, cotangent) -> Tensor:
        cotangent = measurement_vector(cotangent, self.survey.measurement_count, FLOAT_DTYPE)
        return self.vjp_series(self._sigma(conductivity)[None], cotangent[None])[0]

    def vjp_series(self, conductivities, cotangents) -> Tensor:
        """Exact reverse products of a model series (``ert3d.py:744-772``)."""

        s = self._dev
        models = self._series(conductivities)
        cotangents = cotangent_series(cotangents, models.shape[0], self.survey.measurement_count)
        fields, _ = self._fields_series(models)
        gradients = {}
        for batch in newest_batches_first(len(fields), self.series_batch_steps):
            pairs = self._pairs(cotangents[batch].double().numpy())
            rhs = torch.stack([self._combine_receivers(pairs[i], s.rhs_r, s.rp_c) for i in range(len(batch))])
            lam = torch.zeros_like(rhs)
            self._pcg(s, "adjoint", torch.cat([fields[t].values for t in batch]), rhs, lam, s.nc)
            for i, t in enumerate(batch):
                gradients[t] = self._cell_gradient(fields[t].current, s.rp_c, lam[i], s.rp_c, s.nc, robin=True)
        gradient = torch.stack([gradients[t] for t in range(len(fields))]).to(device="cpu", dtype=FLOAT_DTYPE)
        check_finite(gradient, "conductivity gradient")
        return gradient

    def jvp(self, conductivity, delta_conductivity) -> Tensor:
        """Exact forward product (``ert3d.py:774-793``)."""

        s = self._dev
        fields = self._fields(conductivity)
        direction = cell_vector(delta_conductivity, self.mesh.cell_count, torch.float64)
        tangent = self._assemble(s, direction[None], robin=True, fixed=False)
        rhs = torch.zeros((1, self.N, s.rp_c), dtype=torch.float64, device=s.device)
        self._spmm(s, tangent, fields.current[None], rhs, mode=1)
        delta = torch.zeros_like(rhs)
        self._pcg(s, "tangent", fields.values, rhs, delta, s.nc)
        out = self._measurement_resistances(delta)[0].to(device="cpu", dtype=FLOAT_DTYPE)
        check_finite(out, "resistance jvp")
        return out

    def normal_vjp(self, conductivity, cotangent, *, cell_parameter_ids=None, parameter_count=None) -> Tensor:
        cotangent = measurement_vector(cotangent, self.survey.measurement_count, FLOAT_DTYPE)
        return self.normal_vjp_series(self._sigma(conductivity)[None], cotangent[None],
                                      cell_parameter_ids=cell_parameter_ids, parameter_count=parameter_count)[0]

    def normal_vjp_series(self, conductivities, cotangents, *, cell_parameter_ids=None, parameter_count=None) -> Tensor:
        """Transpose of the normal sensitivity (``ert3d.py:814-841``)."""

        s = self._dev
        models = self._series(conductivities)
        cotangents = cotangent_series(cotangents, models.shape[0], self.survey.measurement_count)
        columns = self._columns(cell_parameter_ids, parameter_count)
        fields, _ = self._fields_series(models)
        receivers = self._receivers(fields, models)
        pairs = self._pairs(cotangents.double().numpy())
        grads = []
        for i, f in enumerate(fields):
            lam = self._combine_receivers(pairs[i], receivers[i], s.rp_c)
            grads.append(columns(self._cell_gradient(f.current, s.rp_c, lam, s.rp_c, s.nc, robin=False)))
        out = torch.stack(grads).to(device="cpu", dtype=FLOAT_DTYPE)
        check_finite(out, "normal sensitivity vjp")
        return out

    def normal_jvp(self, conductivity, delta_conductivity) -> Tensor:
        """Normal sensitivity times a direction (``ert3d.py:843-865``)."""

        s = self._dev
        sigma = self._sigma(conductivity)
        fields, _ = self._fields_series(sigma[None])
        (receiver,) = self._receivers(fields, sigma[None])
        direction = cell_vector(delta_conductivity, self.mesh.cell_count, torch.float64)
        tangent = self._assemble(s, direction[None], robin=False, fixed=False)
        y = torch.zeros((1, self.N, s.rp_c), dtype=torch.float64, device=s.device)
        self._spmm(s, tangent, fields[0].current[None], y, mode=1)
        gram = y[0][:, : s.nc].T.contiguous() @ receiver[:, : s.nr].contiguous()
        gram = torch.nn.functional.pad(gram, (0, 1, 0, 1))
        a, b, m, n = s.abmn
        out = (gram[a, m] - gram[b, m] - gram[a, n] + gram[b, n]).to(device="cpu", dtype=FLOAT_DTYPE)
        check_finite(out, "normal sensitivity jvp")
        return out

    def jacobian(self, conductivity, *, batch_size=None, include_robin_boundary_derivative: bool = False,
                 normal_sensitivity: bool = True, jacobian_cell_parameter_ids=None, jacobian_parameter_count=None,
                 return_device: bool = False) -> Tensor:
        """``-(u_M - u_N)^T T_c (u_A - u_B)`` by the fused kernel (``ert3d.py:867-927``),
        ``batch_size`` measurement rows per launch.  An explicit ``batch_size`` bounds the device
        block as in the reference; by default every row goes in one launch unless the (rows,
        cells) FP64 block would exceed ``JACOBIAN_BLOCK_BYTES`` (the 288 GB device holds the
        17 GB block of 2000 rows x 1.1M cells)."""

        del normal_sensitivity
        s = self._dev
        count = self.survey.measurement_count
        if batch_size is None:
            batch_size = jacobian_rows_per_launch(count, self.mesh.cell_count)
        batch_size = int(batch_size)
        if batch_size < 1:
            raise ValueError("batch_size must be >= 1")
        columns = self._columns(jacobian_cell_parameter_ids, jacobian_parameter_count)
        sigma = self._sigma(conductivity)
        fields, _ = self._fields_series(sigma[None])
        (receiver,) = self._receivers(fields, sigma[None])
        all_templates = (s.templates if include_robin_boundary_derivative else s.volume_templates).contiguous()
        precision = 1 if self.sensitivity_precision == "mixed" else 0
        if precision == 0 and s.nc <= 128 and s.nr <= 128:
            # cross-Gram form on the interleaved device-order blocks: per column G = sum_c Phi_R^T
            # T_c Phi_C (receivers x currents; aggregated parameter columns sum their cells inside
            # the kernel), then four reads per row (FP64; the mixed mode keeps the per-output FP32
            # kernel)
            current = fields[0].current.contiguous()
            rec = receiver.contiguous()
            a, b, m, n = s.abmn_i32
            fused = columns.kernel_map

            def launch(start, stop, part):
                native.sensitivity_gram_cross(rec, [0], rec.shape[-1], s.nr, current, [0], current.shape[-1], s.nc,
                                              a[start:stop], b[start:stop], m[start:stop], n[start:stop], s.cell_dofs,
                                              all_templates, None, [1.0], None, part, columns=fused)
            return chunked_rows(count, columns.count, batch_size, launch, None, columns.count, s.device,
                                return_device)
        templates = columns.select(all_templates, 0).contiguous()
        S = s.nc + s.nr
        phi = torch.empty((1, S, self.N), dtype=torch.float64, device=s.device)  # mesh DOF order
        phi[0, : s.nc] = fields[0].current[s.inv, : s.nc].T
        phi[0, s.nc:] = receiver[s.inv, : s.nr].T
        a, b, m, n = (x.copy() for x in s.abmn_host)
        a = np.where(a == s.nc, S, a)
        b = np.where(b == s.nc, S, b)
        m, n = m + s.nc, n + s.nc
        idx = [torch.as_tensor(x.astype(np.int32), device=s.device) for x in (a, b, m, n)]
        dofs = columns.select(s.cell_dofs_mesh, 0).contiguous()

        def launch(start, stop, part):
            native.sensitivity(phi, *(x[start:stop] for x in idx), dofs, templates, None, [1.0], None, part,
                               precision)
        return chunked_rows(count, dofs.shape[0], batch_size, launch, columns.finish(), columns.count, s.device,
                            return_device)

    def jacobian_columnwise(self, conductivity) -> Tensor:
        basis = torch.eye(self.mesh.cell_count, dtype=FLOAT_DTYPE)
        return torch.stack([self.jvp(conductivity, column) for column in basis], dim=1)

    def solve_with_jacobian(self, conductivity, currents=1.0, *, batch_size=None,
                            include_robin_boundary_derivative: bool = False, normal_sensitivity: bool = True,
                            include_fields: bool = False, jacobian_cell_parameter_ids=None,
                            jacobian_parameter_count=None):
        sigma = self._sigma(conductivity)
        response = self._response(self._fields(sigma), currents, include_fields=include_fields)
        jacobian = self.jacobian(sigma, batch_size=batch_size,
                                 include_robin_boundary_derivative=include_robin_boundary_derivative,
                                 normal_sensitivity=normal_sensitivity,
                                 jacobian_cell_parameter_ids=jacobian_cell_parameter_ids,
                                 jacobian_parameter_count=jacobian_parameter_count)
        return response, jacobian

    def prepare(self, conductivity=None, *, include_solver_state: bool = True) -> None:
        self._geometric_factors()
        if include_solver_state:
            self._fields(torch.ones(self.mesh.cell_count, dtype=FLOAT_DTYPE) if conductivity is None else conductivity)

    def close(self) -> None:
        self._field_cache.clear()
        self._solver.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass
