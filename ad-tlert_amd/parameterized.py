"""Consumers of the forward path that map inversion parameters onto forward cells.

* ``log_response_and_jacobian`` (``forward/modeling.py:137-162``): ``log rho_a`` and
  ``d log rho_a / d log rho`` from the operator's resistance Jacobian; with a cell-to-parameter
  map the columns are aggregated inside the Gram sensitivity kernels.
* ``ParameterizedERTForward2p5D`` (``inversion/parameterized.py:28-360``): a forward operator whose
  parameters live on a subset (or coarsening) of the forward cells.  Cells outside the parameter
  domain take their resistivity from a background rule: ``"pygimli_prolongation"`` (neighbour-
  weighted averages, assigned ring by ring inward from the parameter domain), ``"nearest"`` (the
  parameter whose cells' centre is nearest) or ``"fixed_mean"`` (the mean log resistivity).
  ``forward_and_jacobian`` with the prolongation rule, normal sensitivities and no Robin
  derivative is the fused case: one aggregated Jacobian launch whose columns are the parameters
  (background cells dropped), chained with each parameter's conductivity.
* ``normal_log_response_vjp_series`` (``inversion/matrix_free.py:41-100``): rows of
  ``(d log rho_a / d log rho)^T g`` from the operators' ``normal_vjp`` without materializing a
  Jacobian (fused aggregation under the prolongation rule, the projection's transpose otherwise).

The validation rules and messages are the reference's (callers match on them).  All device work
is the operators' (native library); this module is host-side bookkeeping.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch
from scipy.spatial import cKDTree

from .dtypes import FLOAT_DTYPE
from .ert2p5d import ERTForward2p5D
from .ert3d import ERTForward3D
from .mesh import Mesh3D

BACKGROUND_MODES = ("pygimli_prolongation", "nearest", "fixed_mean")


def log_response_and_jacobian(operator, conductivity, *, chain=None, **solve_kwargs):
    """``(log rho_a, d log rho_a / d log rho)`` as numpy arrays.

    With ``rho_a = k R / I``: ``d log rho_a / d log rho_c = -(k / rho_a) dR/dsigma_c sigma_c``;
    ``chain`` is the conductivity of each Jacobian column (the parameters' when the columns are
    aggregated), ``conductivity`` itself by default."""

    response, resistance_jacobian = operator.solve_with_jacobian(conductivity, **solve_kwargs)
    rhoa = torch.as_tensor(response.apparent_resistivity, dtype=torch.float64)
    factors = torch.as_tensor(operator._geometric_factors(), dtype=torch.float64)
    column_sigma = torch.as_tensor(conductivity if chain is None else chain, dtype=torch.float64)
    jacobian = resistance_jacobian.to(torch.float64) * (factors / rhoa)[:, None] * (-column_sigma)[None, :]
    return np.log(rhoa.numpy()), jacobian.numpy()


def as_log_model(model, expected_size: int, log_model: bool, name: str) -> np.ndarray:
    """A flat log-resistivity vector (``inversion/inputs.py:35-53`` rules and messages)."""

    values = np.asarray(model, dtype=float).reshape(-1)
    if values.shape != (expected_size,):
        raise ValueError(f"{name} must have shape ({expected_size},)")
    if not np.all(np.isfinite(values)):
        raise ValueError(f"{name} contains non-finite values")
    if log_model:
        return values.copy()
    if np.any(values <= 0.0):
        raise ValueError(f"{name} must contain positive resistivity values")
    return np.log(values)


def neighbor_pairs(mesh):
    """Pairs of cells sharing an interior edge (2D) or face (3D), with ``|n_z|`` of that
    interface's unit normal (``inversion/regularization.py:97-120``), in first-appearance order
    of the shared entity."""

    nodes = np.asarray(mesh.nodes, dtype=float)
    cells = np.asarray(mesh.cells, dtype=np.int64)
    three_d = isinstance(mesh, Mesh3D)
    if three_d:
        local = np.array([[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]])
    else:
        k = cells.shape[1]
        local = np.stack((np.arange(k), (np.arange(k) + 1) % k), axis=1)
    entities = cells[:, local]  # (C, faces, nodes per face)
    owner = np.repeat(np.arange(cells.shape[0]), local.shape[0])
    keys = np.sort(entities.reshape(-1, local.shape[1]), axis=1)
    _, first, inverse, counts = np.unique(keys, axis=0, return_index=True, return_inverse=True, return_counts=True)
    inverse = inverse.reshape(-1)
    shared = np.flatnonzero(counts[inverse] == 2)
    order = np.lexsort((shared, first[inverse[shared]]))  # by entity first appearance, then owner order
    shared = shared[order]
    pair_owner = owner[shared].reshape(-1, 2)
    points = nodes[entities.reshape(-1, local.shape[1])[shared[::2]]]
    if three_d:
        normal = np.cross(points[:, 1] - points[:, 0], points[:, 2] - points[:, 0])
    else:
        tangent = points[:, 1] - points[:, 0]
        normal = np.stack((-tangent[:, 1], tangent[:, 0]), axis=1)
    length = np.linalg.norm(normal, axis=1)
    keep = length > 0.0
    return [tuple(map(int, p)) for p in pair_owner[keep]], np.abs(normal[keep, -1]) / length[keep]


class ParameterizedERTForward2p5D:
    """Forward operator with a full solve mesh and a smaller parameter set (2.5D or 3D)."""

    def __init__(self, forward, parameter_cell_ids, *, regularization_mesh=None, forward_cell_parameter_ids=None,
                 background_mode: str = "pygimli_prolongation") -> None:
        self.forward_operator = forward
        self.survey = forward.survey
        self.mesh = forward.mesh
        self.regularization_mesh = regularization_mesh
        n_cells = int(forward.mesh.cell_count)
        self.parameter_cell_ids = np.asarray(parameter_cell_ids, dtype=np.int32).reshape(-1)
        reps = self.parameter_cell_ids
        if reps.size == 0:
            raise ValueError("parameter_cell_ids must not be empty")
        if reps.min() < 0 or reps.max() >= n_cells:
            raise ValueError("parameter_cell_ids reference cells outside the forward mesh")
        if np.unique(reps).size != reps.size:
            raise ValueError("parameter_cell_ids must be unique")
        if background_mode not in BACKGROUND_MODES:
            raise ValueError("background_mode must be 'pygimli_prolongation', 'nearest', or 'fixed_mean'")
        self.background_mode = background_mode
        if forward_cell_parameter_ids is None:
            ids = np.full(n_cells, -1, dtype=np.int32)
            ids[reps] = np.arange(reps.size, dtype=np.int32)
        else:
            ids = np.asarray(forward_cell_parameter_ids, dtype=np.int32).reshape(-1)
            if ids.shape != (n_cells,):
                raise ValueError(f"forward_cell_parameter_ids must have one entry per forward mesh cell "
                                 f"({ids.shape} != ({n_cells},))")
            if np.any(ids < -1):
                raise ValueError("forward_cell_parameter_ids may only contain -1 or non-negative parameter ids")
            used = np.unique(ids[ids >= 0])
            if used.size == 0:
                raise ValueError("forward_cell_parameter_ids must contain at least one active parameter cell")
            if not np.array_equal(used, np.arange(used[-1] + 1)):
                raise ValueError("forward_cell_parameter_ids must use contiguous ids starting at 0")
        self.forward_cell_parameter_ids = ids
        self._n_parameters = int(ids.max()) + 1
        if reps.size != self._n_parameters:
            raise ValueError("parameter_cell_ids must contain one representative forward cell per inversion parameter "
                             f"({reps.size} != {self._n_parameters})")
        if not np.array_equal(ids[reps], np.arange(self._n_parameters)):
            raise ValueError("parameter_cell_ids must be ordered representatives of forward_cell_parameter_ids")
        if regularization_mesh is not None and int(regularization_mesh.cell_count) != self._n_parameters:
            raise ValueError("regularization mesh cell count must match inversion parameter count "
                             f"({regularization_mesh.cell_count} != {self._n_parameters})")
        every = np.arange(n_cells, dtype=np.int32)
        self._active_forward_mask = ids >= 0
        self._active_forward_cell_ids = every[self._active_forward_mask]
        self._active_parameter_ids = ids[self._active_forward_cell_ids]
        self.background_cell_ids = every[~self._active_forward_mask]
        self._background_parameter_ids = (self._nearest_parameters() if background_mode == "nearest"
                                          else np.empty(0, dtype=np.int32))
        self._resistivity_prolongation_matrix = (self._prolongation() if background_mode == "pygimli_prolongation"
                                                 else None)
        self._jacobian_projection = self._projection()
        self._torch_maps: dict[str, torch.Tensor] = {}

    @classmethod
    def from_mesh_survey(cls, mesh, survey, parameter_cell_ids, *, regularization_mesh=None,
                         background_mode: str = "pygimli_prolongation", **forward_kwargs):
        ids = forward_kwargs.pop("forward_cell_parameter_ids", None)
        operator = (ERTForward3D if isinstance(mesh, Mesh3D) else ERTForward2p5D).from_mesh_survey(
            mesh, survey, **forward_kwargs)
        return cls(operator, parameter_cell_ids, regularization_mesh=regularization_mesh,
                   forward_cell_parameter_ids=ids, background_mode=background_mode)

    @property
    def cell_count(self) -> int:
        return self._n_parameters

    def close(self) -> None:
        self.forward_operator.close()

    # -- parameter -> forward-cell maps -------------------------------------------------------

    def _centers(self) -> np.ndarray:
        return np.asarray(self.mesh.nodes, dtype=float)[np.asarray(self.mesh.cells)].mean(axis=1)

    def _nearest_parameters(self) -> np.ndarray:
        """Per background cell: the parameter whose cells' mean centre is nearest."""

        if self.background_cell_ids.size == 0:
            return np.empty(0, dtype=np.int32)
        centers = self._centers()
        counts = np.bincount(self._active_parameter_ids, minlength=self.cell_count).astype(float)
        if np.any(counts <= 0.0):
            raise ValueError("each inversion parameter must own at least one forward mesh cell")
        means = np.stack([np.bincount(self._active_parameter_ids, weights=centers[self._active_forward_cell_ids, d],
                                      minlength=self.cell_count) for d in range(centers.shape[1])], axis=1)
        _, nearest = cKDTree(means / counts[:, None]).query(centers[self.background_cell_ids])
        return np.asarray(nearest, dtype=np.int32)

    def _projection(self) -> sp.csr_matrix:
        """``d log rho_forward / d log rho_parameter`` (one-hot on parameter cells; the nearest
        parameter or the mean over all parameters on background cells)."""

        rows, cols, data = [self._active_forward_cell_ids], [self._active_parameter_ids], \
            [np.ones(self._active_forward_cell_ids.size)]
        bg = self.background_cell_ids
        if bg.size and self.background_mode == "nearest":
            rows.append(bg)
            cols.append(self._background_parameter_ids)
            data.append(np.ones(bg.size))
        elif bg.size and self.background_mode == "fixed_mean":
            rows.append(np.repeat(bg, self.cell_count))
            cols.append(np.tile(np.arange(self.cell_count, dtype=np.int32), bg.size))
            data.append(np.full(bg.size * self.cell_count, 1.0 / self.cell_count))
        return sp.coo_matrix((np.concatenate(data), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(int(self.mesh.cell_count), self.cell_count)).tocsr()

    def _prolongation(self) -> sp.csr_matrix:
        """Resistivity prolongation ``rho_forward = P rho_parameter``: one-hot rows on parameter
        cells; background cells, ring by ring inward from the parameter domain (cells in id
        order within a ring), take the average of their already-known neighbours' rows weighted
        by ``sqrt(1 - n_z^2) + 1e-6`` of the shared interface (horizontal neighbours weigh most)."""

        n_cells = int(self.mesh.cell_count)
        rows = np.zeros((n_cells, self.cell_count))
        rows[self._active_forward_cell_ids, self._active_parameter_ids] = 1.0
        if self.background_cell_ids.size == 0:
            return sp.csr_matrix(rows)
        pairs, normal_z = neighbor_pairs(self.mesh)
        pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
        weight = np.sqrt(np.clip(1.0 - normal_z ** 2, 0.0, None)) + 1.0e-6
        # both directions of every pair; per cell in pair order, as neighbour lists are built
        src = np.concatenate((pairs[:, 0], pairs[:, 1]))
        dst = np.concatenate((pairs[:, 1], pairs[:, 0]))
        w = np.concatenate((weight, weight))
        seq = np.concatenate((2 * np.arange(len(pairs)), 2 * np.arange(len(pairs)) + 1))
        order = np.lexsort((seq, src))
        src, dst, w = src[order], dst[order], w[order]
        start = np.searchsorted(src, np.arange(n_cells + 1))
        known = self._active_forward_mask.copy()
        pending = np.sort(self.background_cell_ids)
        while pending.size:
            ring = []
            for cell in pending:
                nb = dst[start[cell]:start[cell + 1]]
                wt = w[start[cell]:start[cell + 1]]
                use = known[nb]
                total = float(wt[use].sum()) if use.any() else 0.0
                if total > 1.0e-8:
                    acc = np.zeros(self.cell_count)
                    for j, ww in zip(nb[use], wt[use]):
                        acc += rows[j] * ww
                    ring.append((cell, acc / total))
            if not ring:
                raise ValueError("could not prolongate background forward cells from active parameter cells")
            for cell, row in ring:
                rows[cell] = row
                known[cell] = True
            done = {cell for cell, _ in ring}
            pending = np.asarray([c for c in pending if c not in done], dtype=np.int64)
        return sp.csr_matrix(rows)

    def _full_log_model(self, log_resistivity) -> np.ndarray:
        return self._full_log_model_and_projection(log_resistivity)[0]

    def _full_log_model_and_projection(self, log_resistivity):
        """Forward-cell log resistivities of a parameter log model, and the Jacobian projection."""

        logs = np.asarray(log_resistivity, dtype=float).reshape(-1)
        if logs.shape != (self.cell_count,):
            raise ValueError(f"log_resistivity must have shape ({self.cell_count},)")
        if self.background_mode == "pygimli_prolongation":
            full = np.asarray(self._resistivity_prolongation_matrix @ np.exp(logs), dtype=float).reshape(-1)
            if np.any(full <= 0.0) or not np.all(np.isfinite(full)):
                raise ValueError("prolongated forward resistivity contains non-positive or non-finite values")
            return np.log(full), self._jacobian_projection
        full = np.empty(int(self.mesh.cell_count))
        full[self._active_forward_cell_ids] = logs[self._active_parameter_ids]
        if self.background_cell_ids.size:
            full[self.background_cell_ids] = (logs[self._background_parameter_ids] if self.background_mode == "nearest"
                                              else float(np.mean(logs)))
        return full, self._jacobian_projection

    def log_model_to_full(self, log_parameters: torch.Tensor) -> torch.Tensor:
        """Differentiable ``(..., n_parameters) -> (..., n_forward_cells)`` log-resistivity map
        (float64, on the input's device)."""

        exponential = self.background_mode == "pygimli_prolongation"
        device = log_parameters.device
        key = str(device)
        if key not in self._torch_maps:
            m = (self._resistivity_prolongation_matrix if exponential else self._jacobian_projection).tocoo()
            self._torch_maps[key] = torch.sparse_coo_tensor(np.vstack((m.row, m.col)), m.data, m.shape,
                                                            dtype=torch.float64, device=device).coalesce()
        flat = log_parameters.to(torch.float64).reshape(-1, log_parameters.shape[-1]).T
        full = torch.sparse.mm(self._torch_maps[key], torch.exp(flat) if exponential else flat).T
        if exponential:
            full = torch.log(full)
        return full.reshape(*log_parameters.shape[:-1], full.shape[-1])

    # -- forward products ----------------------------------------------------------------------

    def forward_and_jacobian(self, resistivity_model, log_transform: bool = True, *,
                             include_robin_boundary_derivative=None, normal_sensitivity=None):
        """``(log rho_a, d log rho_a / d log rho_parameter)`` (``inversion/parameterized.py:302-327``)."""

        log_model = as_log_model(resistivity_model, self.cell_count, log_transform, "resistivity_model")
        full_log, projection = self._full_log_model_and_projection(log_model)
        robin = bool(include_robin_boundary_derivative)
        normal = True if normal_sensitivity is None else bool(normal_sensitivity)
        sigma = torch.as_tensor(np.exp(-full_log), dtype=FLOAT_DTYPE)
        if self.background_mode == "pygimli_prolongation" and normal and not robin:
            return log_response_and_jacobian(self.forward_operator, sigma, chain=np.exp(-log_model),
                                             jacobian_cell_parameter_ids=self.forward_cell_parameter_ids,
                                             jacobian_parameter_count=self.cell_count)
        predicted, full_jacobian = log_response_and_jacobian(
            self.forward_operator, sigma, include_robin_boundary_derivative=robin, normal_sensitivity=normal)
        return predicted, np.asarray((projection.T @ full_jacobian.T).T, dtype=float)

    def forward(self, resistivity_model, log_transform: bool = True) -> np.ndarray:
        log_model = as_log_model(resistivity_model, self.cell_count, log_transform, "resistivity_model")
        sigma = torch.as_tensor(np.exp(-self._full_log_model(log_model)), dtype=FLOAT_DTYPE)
        response = np.log(np.asarray(self.forward_operator.apparent_resistivity_values(sigma), dtype=float))
        return response if log_transform else np.exp(response)

    def response(self, resistivity_model) -> np.ndarray:
        return self.forward(resistivity_model, log_transform=False)


def _native_operator(forward):
    if isinstance(forward, (ERTForward2p5D, ERTForward3D)):
        return forward
    inner = getattr(forward, "forward_operator", None)
    return inner if isinstance(inner, (ERTForward2p5D, ERTForward3D)) else None


def normal_log_response_vjp_series(forward, log_resistivity, predicted_log, cotangents) -> np.ndarray:
    """Rows of ``(d log rho_a / d log rho)^T g`` per step without materializing Jacobians
    (``inversion/matrix_free.py:41-100``): ``g`` becomes a resistance cotangent ``g k / rho_a``,
    the operator's ``normal_vjp`` gives ``dR/dsigma``, chained by ``-sigma``; the prolongation
    rule aggregates over parameters inside the operator, the other rules project on the host."""

    operator = _native_operator(forward)
    if operator is None:
        raise TypeError("matrix-free normal VJP requires an ERTForward2p5D or ERTForward3D operator")
    logs = np.asarray(log_resistivity, dtype=float)
    predicted = np.asarray(predicted_log, dtype=float)
    cot = np.asarray(cotangents, dtype=float)
    if logs.ndim != 2 or predicted.ndim != 2 or cot.shape != predicted.shape:
        raise ValueError("matrix-free VJP series inputs must be compatible 2D arrays")
    factors = np.asarray(operator._geometric_factors(), dtype=float).reshape(1, -1)
    resistance_cot = cot * factors / np.exp(predicted)
    parameterized = isinstance(forward, ParameterizedERTForward2p5D)
    fused = parameterized and forward.background_mode == "pygimli_prolongation"
    options = dict(cell_parameter_ids=forward.forward_cell_parameter_ids,
                   parameter_count=forward.cell_count) if fused else {}
    rows = []
    for log_model, g in zip(logs, resistance_cot):
        full_log, projection = forward._full_log_model_and_projection(log_model) if parameterized else (log_model, None)
        grad_sigma = np.asarray(operator.normal_vjp(torch.as_tensor(np.exp(-full_log), dtype=FLOAT_DTYPE),
                                                    torch.as_tensor(g, dtype=FLOAT_DTYPE), **options),
                                dtype=float).reshape(-1)
        if fused:
            rows.append(-np.exp(-log_model) * grad_sigma)
        else:
            per_cell = -np.exp(-full_log) * grad_sigma
            rows.append(per_cell if projection is None else np.asarray(projection.T @ per_cell, dtype=float).reshape(-1))
    return np.vstack(rows)
