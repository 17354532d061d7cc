"""2.5D multi-wavenumber ERT forward operator on one MI355X.

Drop-in for ``adtlert.forward.ERTForward2p5D`` (``forward/ert2p5d.py:108-1121``).  Every
cosine-transform wavenumber ``w`` and source electrode ``s`` solves

    A_w(sigma) u_ws = A_w(1) u_p,ws

with the reference's secondary-field split ``u = u_s + rho_src u_p`` used as the initial guess
``x0 = rho_src u_p`` of an FP64 CG-family solver (mathematically the same system the reference
factorizes with cuDSS).  Assembly, solves, electrode potentials, the fused sensitivity and the
adjoint gradient run in the native gfx950 library; results return to the host in
``FLOAT_DTYPE`` like the reference.

On a flat surface the fields live on the input mesh with the analytic primary field.  On
terrain (``ert2p5d.py:170-190, 349-436, 662-671``) they live on an auxiliary space built by
``terrain.build`` (H2-refined quad strips or refined P1 triangles), the primary field is the
unit-conductivity potential of a serendipity/P2 space solved on the device, geometric factors
may be numerical, and cell quantities are summed back to the input cells.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch
from scipy.special import k0 as bessel_k0

from . import fem, formats, native, terrain
from .ordering import dof_order, inverse
from .dtypes import FLOAT_DTYPE
from .field_cache import FieldCache, digest16, newest_batches_first
from .inputs import (cell_vector, check_finite, conductivity_model, conductivity_series, cotangent_series,
                     measurement_vector, option, parameter_targets)
from .mesh import Mesh
from .quadrature import build_inverse_cosine_weights, survey_wavenumber_bounds
from .solver import BatchedSolver, choose_lanes
from .survey import Survey

Tensor = torch.Tensor
PYGIMLI_ELECTRODE_RADIUS_FRACTION = 1.0 / 6.0
"""pyGIMLi evaluates the 2.5D primary field at an electrode node with K0 at this fraction of the
distance to the nearest neighbouring node (``ert2p5d.py:61``)."""


@dataclass(frozen=True)
class ForwardResponse:
    """Result of a 2.5D forward solve (``ert2p5d.py:66-76``)."""

    apparent_resistivity: Tensor
    resistance: Tensor
    electrode_potentials: Tensor
    integrated_potentials: Tensor
    wavenumbers: Tensor
    weights: Tensor


def _halfspace_primary(nodes: np.ndarray, source: np.ndarray, wavenumber: float, surface: float) -> np.ndarray:
    """Unit-resistivity half-space primary field, image source when buried (``ert2p5d.py:78-94``)."""

    distance = np.linalg.norm(nodes - source, axis=1)
    values = np.zeros(nodes.shape[0])
    if abs(source[1] - surface) <= 1e-8:
        ok = distance > 1e-12
        values[ok] = bessel_k0(distance[ok] * wavenumber) / np.pi
        return values
    mirrored = np.linalg.norm(nodes - (source[0], 2.0 * surface - source[1]), axis=1)
    ok = (distance > 1e-12) & (mirrored > 1e-12)
    values[ok] = (bessel_k0(distance[ok] * wavenumber) + bessel_k0(mirrored[ok] * wavenumber)) / (2.0 * np.pi)
    return values


@dataclass(frozen=True)
class SparsePattern:
    """CSR structure shared by every assembled operator (``discretization.py:36-125``)."""

    shape: tuple[int, int]
    volume_inverse: np.ndarray
    boundary_inverse: np.ndarray
    indptr: np.ndarray
    indices: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.indices.size)


def sparse_pattern(cell_dofs: np.ndarray, boundary_dofs: np.ndarray, n: int) -> SparsePattern:
    """Unique (row, col) pairs of all local entries, in lexicographic order, and each entry's slot."""

    def pairs(conn):
        w = conn.shape[1]
        return np.repeat(conn, w, axis=1).reshape(-1), np.tile(conn, (1, w)).reshape(-1)

    vr, vc = pairs(cell_dofs)
    br, bc = pairs(boundary_dofs) if boundary_dofs.size else (np.zeros(0, np.int64), np.zeros(0, np.int64))
    keys = np.concatenate((vr * n + vc, br * n + bc))
    unique, inverse = np.unique(keys, return_inverse=True)
    rows = unique // n
    indptr = np.concatenate(([0], np.cumsum(np.bincount(rows, minlength=n)))).astype(np.int32)
    return SparsePattern((n, n), inverse[: vr.size], inverse[vr.size:], indptr, (unique % n).astype(np.int32))


def slot_map(inverse: np.ndarray, nnz: int) -> tuple[np.ndarray, np.ndarray]:
    """CSR-slot -> ordered local-entry list (the transpose of the scatter map); entries < 0 dropped."""

    valid = np.flatnonzero(inverse >= 0)
    order = valid[np.argsort(inverse[valid], kind="stable")]
    ptr = np.concatenate(([0], np.cumsum(np.bincount(inverse[valid], minlength=nnz)))).astype(np.int32)
    return ptr, order.astype(np.int32)


@dataclass(frozen=True)
class Discretization:
    """Finite-element space of the flat native path (``discretization.py:128-156, 485-575``)."""

    name: str
    geometry: Mesh
    parent_cell_ids: Tensor
    dof_nodes: Tensor
    cell_dofs: Tensor
    boundary_dofs: Tensor
    stiffness: Tensor
    mass: Tensor
    boundary_mass: Tensor
    boundary_geometries: Tensor
    electrode_matrix: Tensor
    pattern: SparsePattern
    spd: bool

    @property
    def dof_count(self) -> int:
        return int(self.dof_nodes.shape[0])

    @property
    def boundary_cells(self) -> Tensor:
        return self.geometry.boundary_edge_cells.long()


def build_native_discretization(mesh: Mesh, survey: Survey, wavenumbers: np.ndarray, quadrature_order: int = 2):
    """Input-mesh P1-triangle or Q1-quad discretization (``discretization.py:519-575``, ``auxiliary=False``)."""

    if quadrature_order not in (1, 2):
        raise ValueError(f"unsupported triangle quadrature order: {quadrature_order}")
    nodes = np.asarray(mesh.nodes, dtype=float)
    cells = np.asarray(mesh.cells, dtype=np.int64)
    bdofs = np.asarray(mesh.boundary_edges, dtype=np.int64)
    if mesh.is_triangle_mesh:
        stiffness, mass = fem.p1_triangle(nodes, cells)
        if quadrature_order == 1:  # one-point rule: lumped-at-centroid mass
            area = np.asarray(mesh.cell_areas, dtype=float)
            mass = np.broadcast_to(area[:, None, None] / 9.0, mass.shape).copy()
    else:
        stiffness, mass = fem.q1_quad(nodes, cells)
    bmass = fem.edge_mass(np.asarray(mesh.boundary_edge_lengths, dtype=float))
    center = np.mean(np.asarray(survey.electrode_positions, dtype=float), axis=0)
    alpha = np.stack([fem.robin_2p5d(mesh.boundary_edge_centers, mesh.boundary_edge_normals, mesh.surface_edge_mask,
                                     float(mesh.surface_reference_level), center, float(k)) for k in wavenumbers])
    cell_ids, weights = mesh.locate_points(survey.electrode_positions)
    n = nodes.shape[0]
    electrodes = np.zeros((survey.electrode_count, n))
    k = cells.shape[1]
    np.add.at(electrodes, (np.repeat(np.arange(survey.electrode_count), k), cells[np.asarray(cell_ids)].reshape(-1)),
              np.asarray(weights, dtype=float).reshape(-1))
    return Discretization(
        name="native", geometry=mesh, parent_cell_ids=torch.arange(mesh.cell_count), dof_nodes=mesh.nodes,
        cell_dofs=torch.as_tensor(cells), boundary_dofs=torch.as_tensor(bdofs),
        stiffness=torch.as_tensor(stiffness), mass=torch.as_tensor(mass), boundary_mass=torch.as_tensor(bmass),
        boundary_geometries=torch.as_tensor(alpha), electrode_matrix=torch.as_tensor(electrodes),
        pattern=sparse_pattern(cells, bdofs, n), spd=True)


def terrain_discretization(arrays: dict) -> Discretization:
    """A :class:`Discretization` from the arrays of :func:`terrain.build` (auxiliary: not flagged SPD)."""

    cells = np.asarray(arrays["cell_dofs"], dtype=np.int64)
    bdofs = np.asarray(arrays["boundary_dofs"], dtype=np.int64)
    n = int(arrays["dof_nodes"].shape[0])
    return Discretization(
        name=arrays["name"], geometry=arrays["geometry"], parent_cell_ids=torch.as_tensor(arrays["parents"]),
        dof_nodes=torch.as_tensor(arrays["dof_nodes"]), cell_dofs=torch.as_tensor(cells),
        boundary_dofs=torch.as_tensor(bdofs), stiffness=torch.as_tensor(arrays["stiffness"]),
        mass=torch.as_tensor(arrays["mass"]), boundary_mass=torch.as_tensor(arrays["boundary_mass"]),
        boundary_geometries=torch.as_tensor(arrays["alpha"]), electrode_matrix=torch.as_tensor(arrays["electrodes"]),
        pattern=sparse_pattern(cells, bdofs, n), spd=False)


def _solve_unit_potentials(op, d: Discretization) -> np.ndarray:
    """``(W, E, dofs)`` solutions of ``A_w(1) U = E^T`` on discretization ``d`` (one system per
    wavenumber, the electrode rows as shared right-hand sides), assembled and solved on the device."""

    dev = native.device()
    f64 = dict(dtype=torch.float64, device=dev)
    pattern = d.pattern
    n, nnz = d.dof_count, pattern.nnz
    W, E = int(op.wavenumbers.shape[0]), op.survey.electrode_count
    stiff = torch.as_tensor(d.stiffness, **f64)
    mass = torch.as_tensor(d.mass, **f64)
    bmass = torch.as_tensor(d.boundary_mass, **f64)
    parents = d.parent_cell_ids.numpy()
    sptr, sent = slot_map(np.concatenate((pattern.volume_inverse, pattern.boundary_inverse)), nnz)
    k_local, k_bnd = stiff.shape[1], bmass.shape[1]
    amap = native.AssemblyMap(sptr, sent, pattern.volume_inverse.size, k_local * k_local, parents, kb=k_bnd * k_bnd,
                              bnd_col=parents[d.boundary_cells.numpy()], device=dev)
    k = op.wavenumbers.double().numpy()
    values = torch.empty((W, nnz), **f64)
    native.assemble(amap, torch.ones((1, op.mesh.cell_count), **f64), amap.entries(stiff, bmass), values, n_waves=W,
                    emass=amap.entries(mass), k2=torch.as_tensor(k * k, **f64),
                    balpha=torch.as_tensor(d.boundary_geometries, **f64).contiguous(), with_boundary=True)
    lanes = op._solver_lanes(E)
    rp = -(-E // lanes) * lanes
    b = torch.zeros((n, rp), **f64)
    b[:, :E] = torch.as_tensor(d.electrode_matrix.double().numpy().T, **f64)
    x = torch.zeros((W, n, rp), **f64)
    zeros = np.zeros(W, np.int64)
    plan = op._solver.plan((d.name, "unit_potentials"), np.full(W, n, dtype=np.int32), E,
                           pattern_offsets=(zeros, zeros, np.arange(W, dtype=np.int64) * nnz), b_off=zeros,
                           host_pattern=(pattern.indptr, pattern.indices))
    if plan.rp != rp:
        raise RuntimeError(f"interleaved width {rp} does not match the plan's {plan.rp}")
    indptr = torch.as_tensor(pattern.indptr, dtype=torch.int32, device=dev)
    indices = torch.as_tensor(pattern.indices, dtype=torch.int32, device=dev)
    op._solver.solve(plan, indptr, indices, values.contiguous(), b.reshape(-1), x.reshape(-1))
    out = x[:, :, :E].permute(0, 2, 1).contiguous().cpu().numpy()
    check_finite(torch.as_tensor(out), "unit potentials")
    return out


def electrode_stencil(matrix: np.ndarray):
    """Dense (E, N) interpolation rows -> (E, kmax) node ids and weights (zero-weight padding)."""

    rows = [np.flatnonzero(r) for r in matrix]
    kmax = max(1, max(len(r) for r in rows))
    nodes = np.zeros((matrix.shape[0], kmax), dtype=np.int32)
    weights = np.zeros((matrix.shape[0], kmax))
    for i, r in enumerate(rows):
        nodes[i, : len(r)] = r
        weights[i, : len(r)] = matrix[i, r]
    return nodes, weights


def pair_scatter(slots: np.ndarray, count: int, normal: np.ndarray, reciprocal: np.ndarray | None = None):
    """Transpose of the transfer-resistance map onto ``(Q, E, E)`` electrode-potential cotangents
    (``ert2p5d.py:591-629``).  Sequential host scatter: deterministic."""

    width = count + 1
    normal = np.atleast_2d(np.asarray(normal, dtype=np.float64))
    q = normal.shape[0]
    a, b, m, n = (slots[:, i] for i in range(4))
    quads = [((a, b, m, n), normal)]
    if reciprocal is not None:
        quads.append(((m, n, a, b), np.atleast_2d(np.asarray(reciprocal, dtype=np.float64))))
    pairs = np.zeros((q, width * width))
    rows = np.arange(q)[:, None]
    for (src, snk, pos, neg), g in quads:
        for idx, sign in ((src * width + pos, 1.0), (src * width + neg, -1.0), (snk * width + pos, -1.0),
                          (snk * width + neg, 1.0)):
            np.add.at(pairs, (np.broadcast_to(rows, g.shape), np.broadcast_to(idx, g.shape)), sign * g)
    return pairs.reshape(q, width, width)[:, :count, :count]


JACOBIAN_BLOCK_BYTES = 32 << 30
"""Device bytes of one sensitivity launch's (rows, cells) FP64 output by default (an eighth of
the 288 GB of HBM)."""


def jacobian_rows_per_launch(rows: int, cells: int) -> int:
    """Every row in one launch unless the FP64 block would exceed ``JACOBIAN_BLOCK_BYTES``."""

    return max(1, min(rows, JACOBIAN_BLOCK_BYTES // max(8 * cells, 1)))


def chunked_rows(count: int, width: int, batch_size: int, launch, finish, columns: int, device,
                 return_device: bool) -> Tensor:
    """Run ``launch(start, stop, part)`` for consecutive chunks of ``batch_size`` measurement
    rows into a device ``(rows, width)`` block, map each block to its ``columns`` output columns
    with ``finish`` (None: the launch already writes them), and gather the rows on the host in
    ``FLOAT_DTYPE`` (or on the device with ``return_device``).  Without ``finish`` the kernel
    writes straight into the result block."""

    f64 = dict(dtype=torch.float64, device=device)
    done = (lambda part: part) if finish is None else finish
    if count <= batch_size:
        part = torch.empty((count, width), **f64)
        launch(0, count, part)
        result = done(part)
        if return_device:
            return result
        out = result.to(device="cpu", dtype=FLOAT_DTYPE)
    elif return_device:
        result = torch.empty((count, columns), **f64)
        scratch = None if finish is None else torch.empty((batch_size, width), **f64)
        for start in range(0, count, batch_size):
            stop = min(count, start + batch_size)
            if scratch is None:
                launch(start, stop, result[start:stop])
            else:
                launch(start, stop, scratch[: stop - start])
                result[start:stop] = done(scratch[: stop - start])
        return result
    else:
        out = torch.empty((count, columns), dtype=FLOAT_DTYPE)
        part = torch.empty((batch_size, width), **f64)
        for start in range(0, count, batch_size):
            stop = min(count, start + batch_size)
            launch(start, stop, part[: stop - start])
            out[start:stop] = done(part[: stop - start]).to(device="cpu", dtype=FLOAT_DTYPE)
    check_finite(out, "jacobian")
    return out


@dataclass
class _Fields:
    """Assembled values ``(W, nnz)`` and interleaved total fields ``(W, N, Rp)`` of one model,
    in the device DOF numbering; ``inv`` maps mesh nodes to device rows."""

    values: Tensor
    blocks: Tensor
    count: int
    inv: Tensor

    @property
    def rows(self) -> Tensor:
        """Field rows ``(W, E, N)`` in the device numbering (kernel input)."""

        return self.blocks[..., : self.count].transpose(1, 2).contiguous()

    @property
    def total(self) -> Tensor:
        """Field rows ``(W, E, N)`` in mesh node order, as the reference stores them."""

        return self.blocks[:, self.inv, : self.count].transpose(1, 2).contiguous()

    def on(self, device) -> Tensor:
        return self.total.to(device)


class _Columns:
    """Cell -> output-column aggregation (``kernels.py:88-113``) on the device."""

    def __init__(self, targets: np.ndarray, count: int, device):
        self.count = int(count)
        active = np.flatnonzero(targets >= 0)
        self.identity = active.size == targets.size and count == targets.size and np.array_equal(targets, np.arange(count))
        self.active = torch.as_tensor(active, device=device)
        order = np.argsort(targets[active], kind="stable")
        ptr = np.concatenate(([0], np.cumsum(np.bincount(targets[active], minlength=count)))).astype(np.int32)
        self.ptr = torch.as_tensor(ptr, device=device)
        self.idx = torch.as_tensor(order.astype(np.int32), device=device)
        # the sensitivity kernels' fused form: the original cell ids of every column, in order
        self.cells = torch.as_tensor(active[order].astype(np.int32), device=device)

    @property
    def kernel_map(self):
        """``(col_ptr, col_cells)`` for the Gram kernels' fused aggregation; None for identity."""

        return None if self.identity else (self.ptr, self.cells)

    def finish(self):
        """Host-side step after a per-cell launch: None when nothing is left to aggregate."""

        return None if self.identity else self.columns

    def select(self, values: Tensor, dim: int) -> Tensor:
        return values if self.identity else values.index_select(dim, self.active)

    def columns(self, active_values: Tensor) -> Tensor:
        if self.identity:
            return active_values
        x = active_values if active_values.ndim == 2 else active_values[None]
        out = torch.empty((x.shape[0], self.count), dtype=torch.float64, device=x.device)
        native.aggregate_columns(x.contiguous(), self.ptr, self.idx, out)
        return out if active_values.ndim == 2 else out[0]

    def __call__(self, values: Tensor) -> Tensor:
        return self.columns(self.select(values, values.ndim - 1).contiguous())


class _Device:
    """Device-resident operator data of one discretization, in a Hilbert-curve DOF numbering
    (``perm`` maps device rows to mesh nodes, ``inv`` mesh nodes to device rows)."""

    def __init__(self, op: "ERTForward2p5D"):
        d = op.discretization
        dev = native.device()
        self.device = dev
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        perm = dof_order(np.asarray(d.dof_nodes, dtype=float))
        inv = inverse(perm)
        self.perm_host, self.inv_host = perm, inv
        self.inv = torch.as_tensor(inv, dtype=torch.long, device=dev)
        cells = inv[d.cell_dofs.numpy()]
        bdofs = inv[d.boundary_dofs.numpy()]
        pattern = sparse_pattern(cells, bdofs, d.dof_count)
        self.pattern = pattern
        self.indptr = torch.as_tensor(pattern.indptr, **i32)
        self.indices = torch.as_tensor(pattern.indices, **i32)
        nnz = pattern.nnz
        self.nnz = nnz
        inverse_map = np.concatenate((pattern.volume_inverse, pattern.boundary_inverse))
        sptr, sent = slot_map(inverse_map, nnz)
        self.stiff = torch.as_tensor(d.stiffness, **f64).contiguous()
        self.mass = torch.as_tensor(d.mass, **f64).contiguous()
        self.bmass = torch.as_tensor(d.boundary_mass, **f64).contiguous()
        self.balpha = torch.as_tensor(d.boundary_geometries, **f64).contiguous()
        k = op.wavenumbers.double().numpy()
        self.k2_host = k * k
        self.q_host = op.weights.double().numpy()
        self.k2 = torch.as_tensor(self.k2_host, **f64)
        parents = d.parent_cell_ids.numpy()
        self.cell_col = torch.as_tensor(parents, **i32)
        self.bnd_col = torch.as_tensor(parents[d.boundary_cells.numpy()], **i32)
        # slot-ordered scatter map with the template values of every entry (volume entries first)
        k_local, k_bnd = self.stiff.shape[1], self.bmass.shape[1]
        self.amap = native.AssemblyMap(sptr, sent, pattern.volume_inverse.size, k_local * k_local, parents,
                                       kb=k_bnd * k_bnd, bnd_col=parents[d.boundary_cells.numpy()], device=dev)
        self.estiff = self.amap.entries(self.stiff, self.bmass)
        self.emass = self.amap.entries(self.mass)
        self.cell_dofs = torch.as_tensor(cells, **i32).contiguous()
        # the Jacobian kernel walks cells in mesh order and reads fields in mesh node order
        self.cell_dofs_mesh = torch.as_tensor(d.cell_dofs.numpy(), **i32).contiguous()
        self.bnd_dofs = torch.as_tensor(bdofs, **i32).contiguous()
        bcells = d.boundary_cells.numpy()
        order = np.argsort(bcells, kind="stable")
        cptr = np.concatenate(([0], np.cumsum(np.bincount(bcells, minlength=d.cell_dofs.shape[0]))))
        self.cell_bnd = torch.as_tensor(np.concatenate((cptr, order)).astype(np.int32), **i32)
        emat = d.electrode_matrix.double().numpy()[:, perm]
        self.electrodes_t = torch.as_tensor(emat.T.copy(), **f64)  # (N, E)
        en, ew = electrode_stencil(emat)
        self.enode = torch.as_tensor(en, **i32)
        self.eweight = torch.as_tensor(ew, **f64)
        slots = op.survey.measurement_slots().numpy()
        self.slots_host = slots
        self.a, self.b, self.m, self.n = (torch.as_tensor(slots[:, i], **i32) for i in range(4))
        self.along = [torch.as_tensor(slots[:, i], dtype=torch.long, device=dev) for i in range(4)]
        self.N, self.E, self.W = d.dof_count, op.survey.electrode_count, int(op.wavenumbers.shape[0])
        self.lanes = native_lanes = op._solver_lanes(self.E)
        self.rp = -(-self.E // native_lanes) * native_lanes
        # unit primary (W, N, Rp) and the reference right-hand side A(1) u_p
        primary = op._unit_primary()
        self.primary = torch.zeros((self.W, self.N, self.rp), **f64)
        self.primary[:, :, : self.E] = torch.as_tensor(primary[:, :, perm], **f64).transpose(1, 2)
        ones = torch.ones((1, d.cell_dofs.shape[0]), **f64)
        unit_values = op._assemble_device(self, ones)
        self.rhs = torch.empty_like(self.primary)
        op._spmm(self, unit_values, self.primary, self.rhs, self.W)
        # source resistivities: mesh cells around each source node (padded table), else the source cell
        cells = np.asarray(op.mesh.cells)
        n_cells = cells.shape[0]
        lists = [np.flatnonzero(np.any(cells == node, axis=1)) if node >= 0 else np.zeros(0, np.int64)
                 for node in op.source_node_ids.tolist()]
        width = max(1, max(len(x) for x in lists))
        table = np.full((self.E, width), n_cells, dtype=np.int64)
        for e, x in enumerate(lists):
            table[e, : len(x)] = x
        lengths = np.array([len(x) for x in lists])
        self.src_table = torch.as_tensor(table, device=dev)
        self.src_count = torch.as_tensor(np.maximum(lengths, 1), **f64)
        self.src_has_node = torch.as_tensor(lengths > 0, device=dev)
        self.src_fallback = torch.as_tensor(op.source_cell_ids.numpy().astype(np.int64), device=dev)


class ERTForward2p5D:
    """2.5D ERT forward operator on the GPU (interface of ``forward/ert2p5d.py:108-1121``)."""

    def __init__(self, mesh, survey, wavenumbers, weights, discretization, source_cell_ids, source_node_ids,
                 numerical_h2_refined=True, numerical_p2_refined=True, topographic_geometric_factor_mode="analytic",
                 terrain_cache_dir=None, normal_field_cache_max_entries=8, series_batch_steps=8,
                 solver_tolerance=1e-10, sensitivity_precision="float64", preconditioner="amg",
                 primary_potential_discretization=None, geometric_discretization=None):
        self.mesh, self.survey = mesh, survey
        self.wavenumbers, self.weights = wavenumbers, weights
        self.discretization = discretization
        self.primary_potential_discretization = primary_potential_discretization
        self.geometric_discretization = geometric_discretization
        self.source_cell_ids, self.source_node_ids = source_cell_ids, source_node_ids
        self.numerical_h2_refined, self.numerical_p2_refined = numerical_h2_refined, numerical_p2_refined
        self.topographic_geometric_factor_mode = topographic_geometric_factor_mode
        self.terrain_cache_dir = formats.terrain_cache_dir(terrain_cache_dir)
        self.normal_field_cache_max_entries = int(normal_field_cache_max_entries)
        self.series_batch_steps = int(series_batch_steps)
        self.sensitivity_precision = option(sensitivity_precision, ("float64", "mixed"), "sensitivity_precision")
        self._solver = BatchedSolver(tolerance=float(solver_tolerance), preconditioner=option(
            preconditioner, ("amg", "jacobi", "learned", "ic0"), "preconditioner"))
        self._cache: dict = {}
        self._dev_state: _Device | None = None
        self._field_cache_obj = FieldCache(self.normal_field_cache_max_entries)

    @classmethod
    def from_mesh_survey(cls, mesh: Mesh, survey: Survey, quadrature_order: int = 2, numerical_h2_refined: bool = True,
                         numerical_p2_refined: bool = True, topographic_geometric_factor_mode: str = "analytic",
                         terrain_cache_dir=None, normal_field_cache_max_entries: int = 8, series_batch_steps: int = 8,
                         *, solver_tolerance: float = 1e-10, sensitivity_precision: str = "float64",
                         wavenumbers=None, weights=None, preconditioner: str = "amg") -> "ERTForward2p5D":
        """Build the operator (``ert2p5d.py:132-216``).  ``wavenumbers``/``weights`` override the
        reference quadrature rule; ``solver_tolerance`` is the relative residual of the iterative
        solves; ``preconditioner`` is ``"amg"`` (multigrid-preconditioned flexible CG),
        ``"learned"`` (the same with the trained damping network) or ``"jacobi"`` (plain Jacobi PCG)."""

        if not torch.cuda.is_available():
            raise RuntimeError("ADTLERT-MI355X requires an AMD GPU with HIP (no device is visible)")
        gf_mode = option(topographic_geometric_factor_mode, ("analytic", "numerical"), "topographic_geometric_factor_mode")
        if int(normal_field_cache_max_entries) < 0:
            raise ValueError("normal_field_cache_max_entries must be non-negative")
        if int(series_batch_steps) < 1:
            raise ValueError("series_batch_steps must be >= 1")
        if wavenumbers is None:
            k, q = build_inverse_cosine_weights(*survey_wavenumber_bounds(survey))
        else:
            k, q = torch.as_tensor(wavenumbers, dtype=FLOAT_DTYPE), torch.as_tensor(weights, dtype=FLOAT_DTYPE)
        potential = geometric = None
        if mesh.is_flat_surface:
            disc = build_native_discretization(mesh, survey, k.double().numpy(), quadrature_order)
        else:
            # terrain (``ert2p5d.py:170-190``): auxiliary field discretization, serendipity/P2
            # primary-potential discretization, and the geometric one for numerical factors
            def build(name, **options):
                return terrain_discretization(terrain.build(name, mesh, survey.electrode_positions, k.double().numpy(),
                                                            quadrature_order=quadrature_order, **options))

            disc = build("primary", refine=numerical_h2_refined, quadratic=False)
            potential = build("primary_potential", refine=numerical_h2_refined, quadratic=True)
            if gf_mode == "numerical":
                geometric = build("geometric", refine=numerical_h2_refined, quadratic=numerical_p2_refined)
        positions = np.asarray(survey.electrode_positions, dtype=float)
        nodes = np.asarray(mesh.nodes, dtype=float)
        dist = np.linalg.norm(nodes[None, :, :] - positions[:, None, :], axis=2)
        nearest = np.argmin(dist, axis=1)
        source_nodes = np.where(dist[np.arange(len(nearest)), nearest] <= 1e-2, nearest, -1)
        return cls(mesh, survey, k, q, disc, mesh.locate_points(survey.electrode_positions)[0].long(),
                   torch.as_tensor(source_nodes, dtype=torch.long), numerical_h2_refined, numerical_p2_refined, gf_mode,
                   terrain_cache_dir, normal_field_cache_max_entries, series_batch_steps, solver_tolerance,
                   sensitivity_precision, preconditioner, potential, geometric)

    # -- properties ------------------------------------------------------------------------------

    @property
    def use_numerical_primary(self) -> bool:
        return self.primary_potential_discretization is not None

    @property
    def use_numerical_geometric_factors(self) -> bool:
        return self.geometric_discretization is not None

    @property
    def dtype(self) -> torch.dtype:
        """Field precision: float64 on terrain, ``FLOAT_DTYPE`` otherwise (the device works in FP64)."""

        return torch.float64 if self.use_numerical_primary else FLOAT_DTYPE

    @property
    def _field_cache(self) -> FieldCache:
        return self._field_cache_obj

    @property
    def _dev(self) -> _Device:
        if self._dev_state is None:
            self._dev_state = _Device(self)
        return self._dev_state

    def _solver_lanes(self, r: int) -> int:
        return choose_lanes(r)

    # -- host precomputation ---------------------------------------------------------------------

    def _cache_arrays(self, d: Discretization):
        return (self.mesh.nodes, self.mesh.cells, self.mesh.surface_node_ids, self.survey.electrode_positions,
                self.survey.measurements, self.wavenumbers, self.weights, d.parent_cell_ids, d.dof_nodes, d.cell_dofs,
                d.boundary_dofs, d.electrode_matrix, d.pattern.indptr, d.pattern.indices, d.stiffness, d.mass,
                d.boundary_mass, d.boundary_geometries)

    def _unit_primary(self) -> np.ndarray:
        """``(W, E, N)`` unit primary field of the field discretization: the analytic half-space field
        on flat surfaces, else the unit potentials of the primary-potential discretization sampled at
        the field DOFs (``ert2p5d.py:404-436``).  Read from the operator's disk cache when
        ``terrain_cache_dir`` (or ``$ADTLERT_TERRAIN_CACHE_DIR``) names one (``ert2p5d.py:349-402``)."""

        d = self.discretization
        shape = (int(self.wavenumbers.shape[0]), self.survey.electrode_count, d.dof_count)
        options = (self.numerical_h2_refined, self.numerical_p2_refined, self.topographic_geometric_factor_mode)
        cache = formats.UnitPotentialCache(self.terrain_cache_dir)
        if not self.use_numerical_primary:
            arrays = (self.mesh.nodes, self.mesh.cells, self.survey.electrode_positions, self.survey.measurements,
                      self.wavenumbers, self.weights, self.source_node_ids, d.dof_nodes, d.cell_dofs)
            return cache.get("native_analytic_primary", "float64", options, arrays, shape, self._analytic_primary)
        potential = self.primary_potential_discretization

        def project():
            selection = terrain.exact_node_indices(potential.dof_nodes.numpy(), d.dof_nodes.numpy())
            return self._unit_potentials(potential)[..., selection]

        return cache.get("auxiliary_sub_potentials", "float64", options, self._cache_arrays(d), shape, project)

    def _unit_potentials(self, d: Discretization) -> np.ndarray:
        """``(W, E, dofs)`` unit-conductivity potentials of every electrode on the auxiliary
        discretization ``d``: ``A_w(1) U_w = electrode rows`` on the device, one system per
        wavenumber sharing the right-hand sides (``ert2p5d.py:391-402``); disk-cached like the primary."""

        shape = (int(self.wavenumbers.shape[0]), self.survey.electrode_count, d.dof_count)
        options = (self.numerical_h2_refined, self.numerical_p2_refined, self.topographic_geometric_factor_mode)
        cache = formats.UnitPotentialCache(self.terrain_cache_dir)
        key = (d.name, "unit_potentials")
        if key not in self._cache:
            self._cache[key] = cache.get(f"{d.name}_sub_potentials", "float64", options, self._cache_arrays(d), shape,
                                         lambda: _solve_unit_potentials(self, d))
        return self._cache[key]

    def _analytic_primary(self) -> np.ndarray:
        """``(W, E, N)`` unit primary with the electrode-node value (``ert2p5d.py:438-468``)."""

        nodes = np.asarray(self.mesh.nodes, dtype=float)
        cells = np.asarray(self.mesh.cells)
        surface = float(np.max(nodes[:, 1]))
        node_ids = self.source_node_ids.tolist()
        radii = {}
        for node in set(node_ids) - {-1}:
            nb = np.setdiff1d(np.unique(cells[np.any(cells == node, axis=1)]), [node])
            radii[node] = np.min(np.linalg.norm(nodes[nb] - nodes[node], axis=1)) if nb.size else None
        positions = np.asarray(self.survey.electrode_positions, dtype=float)
        out = np.zeros((self.wavenumbers.shape[0], len(positions), nodes.shape[0]))
        for w, k in enumerate(self.wavenumbers.double().tolist()):
            for s, (src, node) in enumerate(zip(positions, node_ids)):
                out[w, s] = _halfspace_primary(nodes, src, k, surface)
                if node >= 0:
                    r = radii[node]
                    out[w, s, node] = 0.0 if r is None else bessel_k0(r * PYGIMLI_ELECTRODE_RADIUS_FRACTION * k) / np.pi
        return out

    # -- device building blocks ------------------------------------------------------------------

    def _assemble_device(self, dev: _Device, sigmas: Tensor, *, robin: bool = True) -> Tensor:
        """``(T*W, nnz)`` CSR values of ``A_w(sigma_t)`` (deterministic slot gather)."""

        sigmas = sigmas.to(device=dev.device, dtype=torch.float64).contiguous()
        values = torch.empty((sigmas.shape[0] * dev.W, dev.nnz), dtype=torch.float64, device=dev.device)
        native.assemble(dev.amap, sigmas, dev.estiff, values, n_waves=dev.W, emass=dev.emass, k2=dev.k2,
                        balpha=dev.balpha, with_boundary=robin)
        return values

    def _spmm(self, dev: _Device, values: Tensor, x: Tensor, y: Tensor, n_sys: int, mode: int = 0) -> None:
        rp = x.shape[-1]
        rows = np.full(n_sys, dev.N, dtype=np.int32)
        zeros = np.zeros(n_sys, dtype=np.int64)
        layout = (rows, zeros, zeros, np.arange(n_sys, dtype=np.int64) * dev.nnz,
                  np.arange(n_sys, dtype=np.int64) * dev.N * rp, dev.lanes if rp == dev.rp else 1, rp)
        native.spmm(layout, dev.indptr, dev.indices, values.contiguous(), x.contiguous(), y, mode)

    def _pcg(self, dev: _Device, key, values: Tensor, rhs: Tensor, x: Tensor, r: int, *, shared_rhs: bool) -> None:
        """Solve ``values`` systems (``n_sys = values.shape[0]``) with ``r`` sources in place on interleaved blocks."""

        n_sys = int(values.shape[0])
        rows = np.full(n_sys, dev.N, dtype=np.int32)
        val_off = np.arange(n_sys, dtype=np.int64) * dev.nnz
        b_off = (np.arange(n_sys) % dev.W).astype(np.int64) * dev.N * x.shape[-1] if shared_rhs else None
        plan = self._solver.plan((key, n_sys), rows, r, pattern_offsets=(np.zeros(n_sys, np.int64),
                                 np.zeros(n_sys, np.int64), val_off), b_off=b_off, lanes=None,
                                 host_pattern=(dev.pattern.indptr, dev.pattern.indices))
        if plan.rp != x.shape[-1]:
            raise RuntimeError(f"interleaved width {x.shape[-1]} does not match the plan's {plan.rp}")
        expected = np.arange(n_sys, dtype=np.int64) * dev.N * plan.rp
        if not np.array_equal(plan.vec_off, expected):
            raise RuntimeError("unexpected PCG vector layout")
        self._solver.solve(plan, dev.indptr, dev.indices, values.contiguous(), rhs.contiguous(), x)

    def _source_resistivities(self, dev: _Device, sigmas: Tensor) -> Tensor:
        """``(T, E)`` geometric-mean resistivity of the cells around each source node, else the
        source cell's resistivity (``ert2p5d.py:480-511``).  Fixed-order padded sums."""

        logs = torch.nn.functional.pad(-torch.log(sigmas), (0, 1))  # padding column: log-resistivity 0
        node_rho = torch.exp(logs[:, dev.src_table].sum(dim=-1) / dev.src_count)
        return torch.where(dev.src_has_node, node_rho, 1.0 / sigmas[:, dev.src_fallback])

    def _solve_fields(self, sigmas: Tensor) -> list[_Fields]:
        dev = self._dev
        sigmas = sigmas.to(device=dev.device, dtype=torch.float64)
        T = sigmas.shape[0]
        values = self._assemble_device(dev, sigmas)
        rho = self._source_resistivities(dev, sigmas)
        x = dev.primary[None] * torch.nn.functional.pad(rho, (0, dev.rp - dev.E))[:, None, None, :]
        x = x.reshape(T * dev.W, dev.N, dev.rp).contiguous()
        self._pcg(dev, "forward", values, dev.rhs, x, dev.E, shared_rhs=True)
        x = x.reshape(T, dev.W, dev.N, dev.rp)
        fields = []
        for t in range(T):
            f = _Fields(values[t * dev.W:(t + 1) * dev.W], x[t], dev.E, dev.inv)
            fields.append(f)
        return fields

    def _fields_series(self, models: Tensor):
        return self._field_cache.solve_series(models, self.series_batch_steps, self._solve_fields)

    def _fields(self, conductivity) -> _Fields:
        return self._fields_series(self._sigma(conductivity)[None])[0][0]

    def normal_field_cache_info(self) -> dict[str, int]:
        return self._field_cache.info()

    def _sigma(self, conductivity) -> Tensor:
        return conductivity_model(conductivity, self.mesh.cell_count)

    def _series(self, conductivities) -> Tensor:
        return conductivity_series(conductivities, self.mesh.cell_count)

    # -- measurement -----------------------------------------------------------------------------

    def _potentials(self, blocks: Tensor) -> Tensor:
        """``(T, E, E)`` integrated electrode potentials ``P[t, source, receiver]`` of ``(T, W, N, Rp)``."""

        dev = self._dev
        blocks = blocks.contiguous()
        T = blocks.shape[0]
        out = torch.empty((T, dev.E, dev.E), dtype=torch.float64, device=dev.device)
        offs = np.arange(T * dev.W, dtype=np.int64) * dev.N * blocks.shape[-1]
        native.electrode_potentials(blocks, offs, T, dev.W, blocks.shape[-1], dev.E, dev.enode, dev.eweight, dev.q_host,
                                    out)
        return out

    def _normal_reciprocal(self, potentials: Tensor):
        """Normal (AB source, MN receiver) and reciprocal transfers (``ert2p5d.py:574-589``)."""

        dev = self._dev
        p = torch.nn.functional.pad(potentials, (0, 1, 0, 1))
        a, b, m, n = dev.along

        def transfer(src, snk, pos, neg):
            return (p[..., src, pos] - p[..., src, neg]) - (p[..., snk, pos] - p[..., snk, neg])

        return transfer(a, b, m, n), transfer(m, n, a, b)

    @staticmethod
    def _combine(normal: Tensor, reciprocal: Tensor) -> Tensor:
        return torch.sign(normal) * torch.sqrt(torch.abs(normal * reciprocal))

    @staticmethod
    def _combine_scale(normal: Tensor, reciprocal: Tensor):
        return torch.sign(reciprocal), torch.clamp_min(torch.sqrt(torch.abs(normal * reciprocal)), 1e-30)

    def _geometric_factors(self) -> Tensor:
        if not self.use_numerical_geometric_factors:
            return self.survey.geometric_factors()
        if "geometric_factors" not in self._cache:
            self._cache["geometric_factors"] = self._numerical_geometric_factors()
        return self._cache["geometric_factors"]

    def _numerical_geometric_factors(self) -> Tensor:
        """``1 / (U_AM - U_AN - U_BM + U_BN)`` from the unit potentials of the geometric
        discretization integrated over wavenumbers (``ert2p5d.py:662-671``)."""

        d = self.geometric_discretization
        integrated = np.tensordot(self.weights.double().numpy(), self._unit_potentials(d), axes=([0], [0]))
        potentials = integrated @ d.electrode_matrix.double().numpy().T
        e = potentials.shape[0]
        padded = np.zeros((e + 1, e + 1))
        padded[:e, :e] = potentials
        a, b, m, n = self.survey.measurement_slots().numpy().T
        sources = padded[a] - padded[b]
        rows = np.arange(a.size)
        return torch.as_tensor(1.0 / (sources[rows, m] - sources[rows, n]), dtype=self.dtype)

    def _measure(self, conductivity):
        fields = self._fields(conductivity)
        potentials = self._potentials(fields.blocks[None])
        check_finite(potentials, "integrated potentials")
        normal, reciprocal = self._normal_reciprocal(potentials[0])
        return fields, potentials[0], normal, reciprocal

    def _integrated(self, fields: _Fields) -> Tensor:
        dev = self._dev
        q = torch.as_tensor(dev.q_host, dtype=torch.float64, device=dev.device)
        return torch.tensordot(q, fields.blocks[:, dev.inv, : dev.E], dims=([0], [0])).T.contiguous()

    def _response(self, fields, potentials, normal, reciprocal, currents) -> ForwardResponse:
        resistance = self._combine(normal, reciprocal).cpu().to(FLOAT_DTYPE)
        current = torch.as_tensor(currents, dtype=FLOAT_DTYPE)
        return ForwardResponse(
            apparent_resistivity=self._geometric_factors() * resistance / current, resistance=resistance,
            electrode_potentials=potentials.cpu().to(FLOAT_DTYPE),
            integrated_potentials=self._integrated(fields).cpu().to(FLOAT_DTYPE),
            wavenumbers=self.wavenumbers, weights=self.weights)

    def solve(self, conductivity, currents=1.0) -> ForwardResponse:
        fields, potentials, normal, reciprocal = self._measure(conductivity)
        return self._response(fields, potentials, normal, reciprocal, currents)

    def resistance(self, conductivity) -> Tensor:
        _, _, normal, reciprocal = self._measure(conductivity)
        return self._combine(normal, reciprocal).cpu().to(FLOAT_DTYPE)

    def apparent_resistivity_values(self, conductivity, currents=1.0) -> Tensor:
        return self._geometric_factors() * self.resistance(conductivity) / torch.as_tensor(currents, dtype=FLOAT_DTYPE)

    def resistance_series(self, conductivities) -> Tensor:
        fields, _ = self._fields_series(self._series(conductivities))
        potentials = self._potentials(torch.stack([f.blocks for f in fields]))
        check_finite(potentials, "integrated potentials")
        return self._combine(*self._normal_reciprocal(potentials)).cpu().to(FLOAT_DTYPE)

    def apparent_resistivity_series(self, conductivities, currents=1.0) -> Tensor:
        current = torch.as_tensor(currents, dtype=FLOAT_DTYPE)
        return self._geometric_factors() * self.resistance_series(conductivities) / current

    def prepare(self, conductivity=None, *, include_solver_state: bool = True) -> None:
        _ = self._dev
        if include_solver_state:
            self._fields(torch.ones(self.mesh.cell_count) if conductivity is None else conductivity)

    # -- derivatives -----------------------------------------------------------------------------

    def _columns(self, cell_parameter_ids=None, parameter_count=None) -> _Columns:
        ids, count = parameter_targets(cell_parameter_ids, parameter_count, self.mesh.cell_count)
        targets = ids[self.discretization.parent_cell_ids.numpy()]
        # the default (every cell its own column) needs no digest of its cell_count ids
        key = ("columns", None if cell_parameter_ids is None else digest16(targets.tobytes()), count)
        if key not in self._cache:
            self._cache[key] = _Columns(targets, count, self._dev.device)
        return self._cache[key]

    def _gradient(self, lam: Tensor, lam_rp: int, lam_off, phi: Tensor, phi_rp: int, phi_off, r: int, *, robin: bool,
                  lam_weights=None) -> Tensor:
        """``-sum_ws lam^T dA_w/dsigma phi`` per cell (device, FP64)."""

        dev = self._dev
        q = np.ones(dev.W) if lam_weights is None else np.asarray(lam_weights, float)
        out = torch.empty(dev.cell_dofs.shape[0], dtype=torch.float64, device=dev.device)
        kwargs = {}
        if robin:
            kwargs = dict(bnd_dofs=dev.bnd_dofs, cell_bnd=dev.cell_bnd, bmass=dev.bmass, balpha=dev.balpha, bws=q)
        native.cell_gradient(lam.contiguous(), lam_rp, lam_off, phi.contiguous(), phi_rp, phi_off, r, dev.cell_dofs,
                             dev.stiff, dev.mass, q, q * dev.k2_host, out, **kwargs)
        return out

    def _adjoint_blocks(self, pairs: np.ndarray) -> Tensor:
        """``(Q, E, E)`` pair cotangents -> ``(W, N, Rp)`` adjoint right-hand sides with
        ``Q*E`` interleaved columns: ``w_w * sum_f pairs[q, s, f] e_f`` (``ert2p5d.py:631-639``)."""

        dev = self._dev
        q_count = pairs.shape[0]
        r = q_count * dev.E
        lanes = self._solver_lanes(r)
        rp = -(-r // lanes) * lanes
        p = torch.as_tensor(pairs, dtype=torch.float64, device=dev.device)  # (Q, s, f)
        cols = p.permute(2, 0, 1).reshape(dev.E, r)  # (f, q*E + s)
        base = dev.electrodes_t @ cols  # (N, Q*E)
        out = torch.zeros((dev.W, dev.N, rp), dtype=torch.float64, device=dev.device)
        q = torch.as_tensor(dev.q_host, dtype=torch.float64, device=dev.device)
        out[:, :, :r] = q[:, None, None] * base[None]
        return out

    def _solve_adjoint(self, values: Tensor, rhs: Tensor, r: int, key) -> Tensor:
        dev = self._dev
        lam = torch.zeros_like(rhs)
        self._pcg(dev, key, values, rhs, lam, r, shared_rhs=False)
        return lam

    def vjp(self, conductivity, cotangent) -> Tensor:
        cotangent = measurement_vector(cotangent, self.survey.measurement_count, FLOAT_DTYPE)
        return self.vjp_series(self._sigma(conductivity)[None], cotangent[None])[0]

    def vjp_series(self, conductivities, cotangents) -> Tensor:
        """Exact reverse products of a model series (``ert2p5d.py:835-871``)."""

        dev = self._dev
        models = self._series(conductivities)
        cotangents = cotangent_series(cotangents, models.shape[0], self.survey.measurement_count)
        fields, _ = self._fields_series(models)
        gradients = {}
        for batch in newest_batches_first(len(fields), self.series_batch_steps):
            potentials = self._potentials(torch.stack([fields[t].blocks for t in batch]))
            normal, reciprocal = self._normal_reciprocal(potentials)
            sign, scale = self._combine_scale(normal, reciprocal)
            common = 0.5 * cotangents[batch].to(dev.device, torch.float64) * sign / scale
            g_n, g_r = (common * reciprocal).cpu().numpy(), (common * normal).cpu().numpy()
            rhs = torch.stack([self._adjoint_blocks(pair_scatter(dev.slots_host, dev.E, g_n[i], g_r[i]))
                               for i in range(len(batch))])
            values = torch.cat([fields[t].values for t in batch])
            rp = rhs.shape[-1]
            lam = self._solve_adjoint(values, rhs.reshape(-1, dev.N, rp), dev.E, ("adjoint", rp))
            lam = lam.reshape(len(batch), dev.W, dev.N, rp)
            for i, t in enumerate(batch):
                offs = np.arange(dev.W, dtype=np.int64) * dev.N * rp
                poffs = np.arange(dev.W, dtype=np.int64) * dev.N * dev.rp
                g = self._gradient(lam[i], rp, offs, fields[t].blocks, dev.rp, poffs, dev.E, robin=True)
                gradients[t] = self._columns()(g).cpu().to(FLOAT_DTYPE)  # discretization cells -> mesh cells
        gradient = torch.stack([gradients[t] for t in range(len(fields))])
        check_finite(gradient, "conductivity gradient")
        return gradient

    def jvp(self, conductivity, delta_conductivity) -> Tensor:
        """Exact forward product (``ert2p5d.py:873-890``)."""

        dev = self._dev
        fields, _, normal, reciprocal = self._measure(conductivity)
        direction = cell_vector(delta_conductivity, self.mesh.cell_count).to(dev.device, torch.float64)
        tangent = self._assemble_device(dev, direction[None])
        rhs = torch.zeros_like(fields.blocks)
        self._spmm(dev, tangent, fields.blocks, rhs, dev.W, mode=1)  # rhs = -A(delta) u
        delta = torch.zeros_like(rhs)
        self._pcg(dev, ("tangent", dev.rp), fields.values, rhs, delta, dev.E, shared_rhs=False)
        dn, dr = self._normal_reciprocal(self._potentials(delta[None])[0])
        sign, scale = self._combine_scale(normal, reciprocal)
        out = (0.5 * sign * (dn * reciprocal + normal * dr) / scale).cpu().to(FLOAT_DTYPE)
        check_finite(out, "tangent fields")
        return out

    def normal_vjp(self, conductivity, cotangent, *, cell_parameter_ids=None, parameter_count=None) -> Tensor:
        """Transpose of the normal sensitivity (``ert2p5d.py:892-912``)."""

        dev = self._dev
        columns = self._columns(cell_parameter_ids, parameter_count)
        fields = self._fields(conductivity)
        weights = measurement_vector(cotangent, self.survey.measurement_count, torch.float64).numpy()
        pairs = pair_scatter(dev.slots_host, dev.E, weights[None])[0]  # (s, f)
        pt = torch.as_tensor(pairs, dtype=torch.float64, device=dev.device)
        # current_w[n, e] = sum_f pairs[f, e] phi_w[n, f]
        current = (fields.blocks[..., : dev.E] @ pt).contiguous()  # (W, N, E)
        offs_c = np.arange(dev.W, dtype=np.int64) * dev.N * dev.E
        offs_p = np.arange(dev.W, dtype=np.int64) * dev.N * dev.rp
        g = self._gradient(fields.blocks, dev.rp, offs_p, current, dev.E, offs_c, dev.E, robin=False,
                           lam_weights=dev.q_host)
        out = columns(g).cpu().to(FLOAT_DTYPE)
        check_finite(out, "normal sensitivity vjp")
        return out

    def normal_vjp_series(self, conductivities, cotangents, *, cell_parameter_ids=None, parameter_count=None) -> Tensor:
        models = self._series(conductivities)
        cotangents = cotangent_series(cotangents, models.shape[0], self.survey.measurement_count)
        return torch.stack([self.normal_vjp(m, c, cell_parameter_ids=cell_parameter_ids, parameter_count=parameter_count)
                            for m, c in zip(models, cotangents)])

    def normal_jvp(self, conductivity, delta_conductivity) -> Tensor:
        """Matrix-free normal sensitivity without the Robin derivative (``ert2p5d.py:914-943``)."""

        dev = self._dev
        fields = self._fields(conductivity)
        direction = cell_vector(delta_conductivity, self.mesh.cell_count).to(dev.device, torch.float64)
        tangent = self._assemble_device(dev, direction[None], robin=False)
        y = torch.zeros_like(fields.blocks)
        self._spmm(dev, tangent, fields.blocks, y, dev.W, mode=1)  # y = -A'(delta) u
        q = torch.as_tensor(dev.q_host, dtype=torch.float64, device=dev.device)
        left = fields.blocks[..., : dev.E].transpose(1, 2).contiguous()  # (W, E, N)
        gram = (q[:, None, None] * (left @ y[..., : dev.E].contiguous())).sum(dim=0)
        gram = torch.nn.functional.pad(gram, (0, 1, 0, 1))
        a, b, m, n = dev.along
        out = (gram[m, a] - gram[m, b] - gram[n, a] + gram[n, b]).cpu().to(FLOAT_DTYPE)
        check_finite(out, "normal sensitivity jvp")
        return out

    def _batch_size(self, batch_size, direct: bool) -> int:
        """Measurement rows per sensitivity launch (``ert2p5d.py:971-982`` semantics for an
        explicit value).  The reference's default of 64 rows on large meshes bounds a consumer
        GPU's memory; on the 288 GB device the default takes every row in one launch unless the
        (rows, cells) FP64 block would exceed ``JACOBIAN_BLOCK_BYTES``.  The adjoint path keeps
        the reference's 8."""

        if batch_size is None:
            if not direct:
                return 8
            return jacobian_rows_per_launch(self.survey.measurement_count, self.mesh.cell_count)
        if batch_size < 1:
            raise ValueError("batch_size must be >= 1")
        return int(batch_size)

    def _normal_jacobian(self, fields: _Fields, batch_size: int, columns: _Columns, *, return_device=False) -> Tensor:
        """``-sum_w w_w (u_M - u_N)^T (K + k_w^2 M) (u_A - u_B)`` by the fused kernel
        (``ert2p5d.py:984-1010``), ``batch_size`` measurement rows per launch: the device holds
        one launch's rows at a time unless ``return_device`` asks for the whole block there.  In
        the Gram form the cell-to-parameter aggregation (``kernels.py:88-113``) runs inside the
        kernel: each parameter column sums its cells' Gram blocks before the output phase."""

        dev = self._dev
        ws, wm = dev.q_host, dev.q_host * dev.k2_host
        precision = 1 if self.sensitivity_precision == "mixed" else 0
        D = self.survey.measurement_count
        if dev.E <= 128 and precision == 0:
            # Gram form on the interleaved device-order blocks: per column G = sum_c Phi^T T_c Phi
            # once per launch, then four gathers per row (FP64; the mixed mode keeps the per-output
            # FP32 kernel)
            blocks = fields.blocks.contiguous()
            vec_off = np.arange(dev.W, dtype=np.int64) * dev.N * blocks.shape[-1]
            fused = columns.kernel_map

            def launch(start, stop, part):
                native.sensitivity_gram(blocks, vec_off, blocks.shape[-1], dev.E, dev.a[start:stop], dev.b[start:stop],
                                        dev.m[start:stop], dev.n[start:stop], dev.cell_dofs, dev.stiff, dev.mass, ws,
                                        wm, part, columns=fused)
            return chunked_rows(D, columns.count, batch_size, launch, None, columns.count, dev.device, return_device)
        phi = fields.total  # (W, E, N), mesh node order (cells are walked in mesh order)
        dofs = columns.select(dev.cell_dofs_mesh, 0).contiguous()
        stiff = columns.select(dev.stiff, 0).contiguous()
        mass = columns.select(dev.mass, 0).contiguous()

        def launch(start, stop, part):
            native.sensitivity(phi, dev.a[start:stop], dev.b[start:stop], dev.m[start:stop], dev.n[start:stop],
                               dofs, stiff, mass, ws, wm, part, precision)
        return chunked_rows(D, dofs.shape[0], batch_size, launch, columns.finish(), columns.count, dev.device,
                            return_device)

    def _adjoint_jacobian(self, fields: _Fields, normal, reciprocal, batch_size: int, robin: bool,
                          normal_sensitivity: bool) -> Tensor:
        """Rows from batched adjoint solves with identity cotangents (``ert2p5d.py:1012-1050``)."""

        dev = self._dev
        count = self.survey.measurement_count
        sign, scale = (t.cpu().numpy() for t in self._combine_scale(normal, reciprocal))
        nrm, rcp = normal.cpu().numpy(), reciprocal.cpu().numpy()
        rows = []
        for start in range(0, count, batch_size):
            stop = min(count, start + batch_size)
            eye = np.zeros((stop - start, count))
            eye[np.arange(stop - start), np.arange(start, stop)] = 1.0
            if normal_sensitivity:
                pairs = pair_scatter(dev.slots_host, dev.E, eye)
            else:
                common = 0.5 * eye * sign[None] / scale[None]
                pairs = pair_scatter(dev.slots_host, dev.E, common * rcp[None], common * nrm[None])
            rhs = self._adjoint_blocks(pairs)
            rp = rhs.shape[-1]
            lam = self._solve_adjoint(fields.values, rhs, (stop - start) * dev.E, ("adjoint-jac", rp))
            offs = np.arange(dev.W, dtype=np.int64) * dev.N * rp
            poffs = np.arange(dev.W, dtype=np.int64) * dev.N * dev.rp
            cells = self._columns()  # discretization cells -> mesh cells
            for i in range(stop - start):
                g = self._gradient(lam, rp, offs + i * dev.E, fields.blocks, dev.rp, poffs, dev.E, robin=robin)
                rows.append(cells(g))
        return torch.stack(rows).cpu().to(FLOAT_DTYPE)

    def solve_with_jacobian(self, conductivity, currents=1.0, *, batch_size=None,
                            include_robin_boundary_derivative: bool = False, normal_sensitivity: bool = True,
                            jacobian_cell_parameter_ids=None, jacobian_parameter_count=None):
        fields, potentials, normal, reciprocal = self._measure(conductivity)
        direct = normal_sensitivity and not include_robin_boundary_derivative
        batch_size = self._batch_size(batch_size, direct)
        if direct:
            columns = self._columns(jacobian_cell_parameter_ids, jacobian_parameter_count)
            jacobian = self._normal_jacobian(fields, batch_size, columns)
        else:
            jacobian = self._adjoint_jacobian(fields, normal, reciprocal, batch_size, include_robin_boundary_derivative,
                                              normal_sensitivity)
        return self._response(fields, potentials, normal, reciprocal, currents), jacobian

    def jacobian(self, conductivity, *, batch_size=None, include_robin_boundary_derivative: bool = False,
                 normal_sensitivity: bool = True) -> Tensor:
        return self.solve_with_jacobian(conductivity, batch_size=batch_size,
                                        include_robin_boundary_derivative=include_robin_boundary_derivative,
                                        normal_sensitivity=normal_sensitivity)[1]

    def jacobian_columnwise(self, conductivity) -> Tensor:
        basis = torch.eye(self.mesh.cell_count, dtype=FLOAT_DTYPE)
        return torch.stack([self.jvp(conductivity, column) for column in basis], dim=1)

    def close(self) -> None:
        """Drop cached fields and release the solver workspaces."""

        self._field_cache.clear()
        self._solver.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass
