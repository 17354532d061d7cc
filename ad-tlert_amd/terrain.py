"""Auxiliary discretizations of the 2.5D terrain path (host-side, built once per operator).

On a mesh whose top surface is not flat the reference replaces the analytic half-space primary
field by a numerical one (``forward/discretization.py:159-575``, ``forward/ert2p5d.py:170-190,
349-436, 662-671``):

* the field discretization ("primary"): terrain-following quadrilateral strips (the input quads,
  or a columnar half-cell triangle mesh completed into quads), each quad split into four through
  its edge midpoints and centre when ``numerical_h2_refined``, bilinear (Q1) elements; meshes that
  are not strips keep their triangles (uniformly refined, P1);
* the primary-potential discretization: the same geometry with eight-node serendipity quads (or
  P2 triangles), integrated in FP64 with the 3 x 3 Gauss rule; its unit-conductivity potentials of
  every electrode, sampled at the field DOFs (which are among its nodes), are the unit primary;
* the geometric discretization (``topographic_geometric_factor_mode="numerical"``): serendipity
  or Q1 on the same geometry; its unit potentials give the geometric factors.

Electrodes on a strip mesh interpolate along the top-surface polyline (linear, or quadratic on
the serendipity edge DOFs); elsewhere by point location.  The Robin coefficients are those of the
flat path on the auxiliary geometry's boundary edges.  These builders run on the host; the
assembly, solves and everything downstream run in the native library like the flat path.
"""

from __future__ import annotations

import numpy as np
import torch

from . import fem
from .mesh import Mesh


def midpoint_builder(nodes):
    """A growing point list and ``midpoint(a, b)``, which adds each edge midpoint once."""

    points = [np.asarray(p, dtype=float) for p in np.asarray(nodes, dtype=float)]
    index: dict[tuple[int, int], int] = {}

    def midpoint(a: int, b: int) -> int:
        key = (a, b) if a < b else (b, a)
        found = index.get(key)
        if found is None:
            found = index[key] = len(points)
            points.append(0.5 * (points[key[0]] + points[key[1]]))
        return found

    return points, midpoint, index


def _refined_surface(surface, midpoint):
    surface = [int(s) for s in np.asarray(surface)]
    out = [surface[0]]
    for a, b in zip(surface[:-1], surface[1:]):
        out += [midpoint(a, b), b]
    return out


def refine_quads(mesh: Mesh):
    """Each quad (c0, c1, c2, c3) split into four through its edge midpoints and centre
    (``discretization.py:325-356``); returns the refined mesh and each child's parent."""

    points, midpoint, _ = midpoint_builder(mesh.nodes)
    cells, parents = [], []
    for parent, (c0, c1, c2, c3) in enumerate(np.asarray(mesh.cells).tolist()):
        m01, m12, m23, m30 = midpoint(c0, c1), midpoint(c1, c2), midpoint(c2, c3), midpoint(c3, c0)
        centre = len(points)
        points.append(0.25 * (points[c0] + points[c1] + points[c2] + points[c3]))
        cells += [[c0, m01, centre, m30], [m01, c1, m12, centre], [centre, m12, c2, m23], [m30, centre, m23, c3]]
        parents += [parent] * 4
    surface = _refined_surface(mesh.surface_node_ids, midpoint)
    return Mesh.from_arrays(np.asarray(points), np.asarray(cells), surface_node_ids=surface), np.asarray(parents)


def refine_triangles(mesh: Mesh):
    """Each triangle split into four through its edge midpoints (``mesh/core.py:326-337, 689-698``)."""

    points, midpoint, _ = midpoint_builder(mesh.nodes)
    cells, parents = [], []
    for parent, (n0, n1, n2) in enumerate(np.asarray(mesh.cells).tolist()):
        m01, m12, m20 = midpoint(n0, n1), midpoint(n1, n2), midpoint(n2, n0)
        cells += [[n0, m01, m20], [n1, m12, m01], [n2, m20, m12], [m01, m12, m20]]
        parents += [parent] * 4
    surface = _refined_surface(mesh.surface_node_ids, midpoint)
    return Mesh.from_arrays(np.asarray(points), np.asarray(cells), surface_node_ids=surface), np.asarray(parents)


def serendipity_topology(mesh: Mesh):
    """Eight-node serendipity DOFs on a quad mesh: corners then the midpoints of edges
    (c0c1, c1c2, c2c3, c3c0); boundary edges as (a, midpoint, b) (``discretization.py:359-383``)."""

    points, midpoint, _ = midpoint_builder(mesh.nodes)
    cells = [[c0, c1, c2, c3, midpoint(c0, c1), midpoint(c1, c2), midpoint(c2, c3), midpoint(c3, c0)]
             for c0, c1, c2, c3 in np.asarray(mesh.cells).tolist()]
    boundary = [[a, midpoint(a, b), b] for a, b in np.asarray(mesh.boundary_edges).tolist()]
    return np.asarray(points), np.asarray(cells, dtype=np.int64), np.asarray(boundary, dtype=np.int64)


def p2_topology(mesh: Mesh):
    """P2 DOFs on a triangle mesh: vertices then the midpoints (01, 12, 20); boundary edges as
    (a, b, midpoint) (``mesh/core.py:543-555``)."""

    points, midpoint, _ = midpoint_builder(mesh.nodes)
    cells = [[n0, n1, n2, midpoint(n0, n1), midpoint(n1, n2), midpoint(n2, n0)]
             for n0, n1, n2 in np.asarray(mesh.cells).tolist()]
    boundary = [[a, b, midpoint(a, b)] for a, b in np.asarray(mesh.boundary_edges).tolist()]
    return np.asarray(points), np.asarray(cells, dtype=np.int64), np.asarray(boundary, dtype=np.int64)


def quad_templates(mesh: Mesh, quadratic: bool):
    """Unit-conductivity stiffness, mass and boundary mass on (bilinear-geometry) quads:
    Q1 with the 2 x 2 Gauss rule, or serendipity with 3 x 3 (``discretization.py:223-267``)."""

    nodes, cells = np.asarray(mesh.nodes, dtype=float), np.asarray(mesh.cells, dtype=np.int64)
    lengths = np.asarray(mesh.boundary_edge_lengths, dtype=float)
    if not quadratic:
        stiffness, mass = fem.q1_quad(nodes, cells)
        return stiffness, mass, fem.edge_mass(lengths)
    stiffness, mass = fem.serendipity_quad(nodes, cells)
    return stiffness, mass, fem.quadratic_edge_mass(lengths)


def triangle_templates(mesh: Mesh, quadratic: bool, quadrature_order: int):
    nodes, cells = np.asarray(mesh.nodes, dtype=float), np.asarray(mesh.cells, dtype=np.int64)
    lengths = np.asarray(mesh.boundary_edge_lengths, dtype=float)
    if quadratic:
        stiffness, mass = fem.p2_triangle(nodes, cells)
        return stiffness, mass, fem.quadratic_edge_mass(lengths)
    stiffness, mass = fem.p1_triangle(nodes, cells)
    if quadrature_order == 1:
        area = np.asarray(mesh.cell_areas, dtype=float)
        mass = np.broadcast_to(area[:, None, None] / 9.0, mass.shape).copy()
    return stiffness, mass, fem.edge_mass(lengths)


def _points_on_segments(points, segments, tol):
    start, edge = segments[:, 0], segments[:, 1] - segments[:, 0]
    length_sq = np.sum(edge * edge, axis=1)
    t = np.sum((points[:, None, :] - start) * edge, axis=2) / np.maximum(length_sq, tol)
    distance = np.linalg.norm(points[:, None, :] - (start + t[:, :, None] * edge), axis=2)
    return (t >= -tol) & (t <= 1.0 + tol) & (distance <= tol) & (length_sq > tol)


def surface_interpolation(mesh: Mesh, boundary_dofs, dof_count: int, points, tol: float = 1e-5) -> np.ndarray:
    """``(E, dofs)`` rows interpolating points on the top-surface polyline: a surface node exactly,
    else the linear (or, with the quadratic boundary DOFs (a, m, b), the quadratic) edge
    interpolant of the segment holding the point (``discretization.py:408-450``)."""

    ids = np.asarray(mesh.surface_node_ids, dtype=np.int64)
    surface = np.asarray(mesh.surface_nodes, dtype=float)
    points = np.asarray(points, dtype=float)
    segments = np.stack((surface[:-1], surface[1:]), axis=1)
    hits = _points_on_segments(points, segments, tol)
    mids = None if boundary_dofs is None else {tuple(sorted((a, b))): m for a, m, b in np.asarray(boundary_dofs).tolist()}
    out = np.zeros((points.shape[0], dof_count))
    for row, point in enumerate(points):
        d = np.linalg.norm(surface - point, axis=1)
        nearest = int(np.argmin(d))
        if d[nearest] <= tol:
            out[row, ids[nearest]] = 1.0
            continue
        found = np.flatnonzero(hits[row])
        if found.size == 0:
            raise ValueError("surface point could not be matched to an auxiliary surface edge")
        s = int(found[0])
        start, stop = segments[s]
        edge = stop - start
        length_sq = float(edge @ edge)
        if length_sq <= tol:
            raise ValueError("degenerate surface edge in auxiliary quadrilateral mesh")
        t = min(max(float((point - start) @ edge / length_sq), 0.0), 1.0)
        a, b = int(ids[s]), int(ids[s + 1])
        if mids is None:
            out[row, a], out[row, b] = 1.0 - t, t
        else:
            out[row, a] = 2.0 * (t - 0.5) * (t - 1.0)
            out[row, mids[tuple(sorted((a, b)))]] = 4.0 * t * (1.0 - t)
            out[row, b] = 2.0 * t * (t - 0.5)
    return out


def point_interpolation(mesh: Mesh, cell_dofs, dof_count: int, points) -> np.ndarray:
    """``(E, dofs)`` rows by point location: the containing cell's linear weights, or the P2
    shape values of its barycentric coordinates (``discretization.py:453-473``)."""

    ids, weights = mesh.locate_points(points)
    weights = np.asarray(weights, dtype=float)
    cell_dofs = np.asarray(cell_dofs, dtype=np.int64)
    if cell_dofs.shape[1] == 6:
        l1, l2, l3 = weights.T
        weights = np.stack((l1 * (2 * l1 - 1), l2 * (2 * l2 - 1), l3 * (2 * l3 - 1), 4 * l1 * l2, 4 * l2 * l3,
                            4 * l3 * l1), -1)
    dofs = cell_dofs[np.asarray(ids)]
    out = np.zeros((dofs.shape[0], dof_count))
    np.add.at(out, (np.repeat(np.arange(dofs.shape[0]), dofs.shape[1]), dofs.reshape(-1)), weights.reshape(-1))
    return out


def _orient_ccw(points, tri):
    p0, p1, p2 = (np.asarray(points[i], dtype=float) for i in tri)
    e1, e2 = p1 - p0, p2 - p0
    return tri if e1[0] * e2[1] - e1[1] * e2[0] > 0.0 else [tri[0], tri[2], tri[1]]


def expand_columnar(mesh: Mesh, tol: float = 1e-6):
    """Complete a terrain-following columnar mesh whose strip quads are single triangles (each
    cell spans two adjacent node columns, two nodes in one and one in the other) by adding the
    missing half of every quad, creating a bottom node where a column is one node short
    (``mesh/core.py:466-540``).  Returns ``(mesh, parents)`` or None for other meshes."""

    if not mesh.is_triangle_mesh:
        return None
    nodes = np.asarray(mesh.nodes, dtype=float)
    cells = np.asarray(mesh.cells, dtype=np.int64)
    if nodes.shape[0] == 0 or cells.shape[0] == 0:
        return None
    columns: list[list[int]] = []
    for node in np.argsort(nodes[:, 0], kind="mergesort").tolist():
        if columns and abs(nodes[node, 0] - nodes[columns[-1][0], 0]) <= tol:
            columns[-1].append(node)
        else:
            columns.append([node])
    sizes = [len(c) for c in columns]
    if len(columns) < 10 or min(sizes) < 2 or max(sizes) - min(sizes) > 1:
        return None
    column_of = np.full(nodes.shape[0], -1, dtype=np.int64)
    row_of = np.full(nodes.shape[0], -1, dtype=np.int64)
    for cid, column in enumerate(columns):
        column.sort(key=lambda n: (-nodes[n, 1], nodes[n, 0]))  # top to bottom
        for row, node in enumerate(column):
            column_of[node], row_of[node] = cid, row
    points = nodes.tolist()
    out, parents = [], []
    for parent, cell in enumerate(cells):
        cols, rows = column_of[cell], row_of[cell]
        unique, counts = np.unique(cols, return_counts=True)
        if unique.size != 2 or unique[1] != unique[0] + 1 or sorted(counts.tolist()) != [1, 2]:
            return None
        pair_col, lone_col = int(unique[np.argmax(counts)]), int(unique[np.argmin(counts)])
        pair = cols == pair_col
        order = np.argsort(rows[pair])
        pair_nodes, pair_rows = cell[pair][order], rows[pair][order]
        if pair_rows[1] != pair_rows[0] + 1:
            return None
        lone_row = int(rows[~pair][0])
        if lone_row not in pair_rows:
            return None
        missing_row = int(pair_rows[1] if lone_row == pair_rows[0] else pair_rows[0])
        lone = columns[lone_col]
        if missing_row >= len(lone):
            if missing_row != len(lone) or len(lone) + 1 != max(sizes):
                return None
            depth = nodes[int(pair_nodes[1]), 1] - nodes[columns[pair_col][0], 1]
            points.append([float(nodes[lone[0], 0]), float(nodes[lone[0], 1] + depth)])
            lone.append(len(points) - 1)
        missing = int(lone[missing_row])
        if missing in cell.tolist():
            return None
        out += [cell.tolist(), _orient_ccw(points, [missing, int(pair_nodes[0]), int(pair_nodes[1])])]
        parents += [parent, parent]
    expanded = Mesh.from_arrays(np.asarray(points), np.asarray(out), surface_node_ids=mesh.surface_node_ids)
    return expanded, np.asarray(parents)


def structured_quads(mesh: Mesh):
    """The quadrilateral strip mesh behind a quad mesh or a columnar half-cell triangle mesh
    (each pair of expanded triangles of one input cell becomes one quad), else None
    (``discretization.py:287-322``)."""

    if mesh.is_quadrilateral_mesh:
        return mesh
    expanded = expand_columnar(mesh)
    if expanded is None:
        return None
    triangles, parents = expanded
    nodes = np.asarray(triangles.nodes, dtype=float)
    pairs: dict[int, list[np.ndarray]] = {}
    for cell, parent in zip(np.asarray(triangles.cells), parents):
        pairs.setdefault(int(parent), []).append(cell)
    quads = []
    for parent in range(mesh.cell_count):
        pair = pairs.get(parent, [])
        ids = np.unique(np.concatenate(pair)) if len(pair) == 2 else np.empty(0, dtype=np.int64)
        if ids.size != 4:
            return None
        order = np.argsort(nodes[ids, 0])
        left, right = ids[order[:2]], ids[order[2:]]
        if nodes[left, 0].max() > nodes[right, 0].min() + 1e-6:
            return None
        quads.append([left[np.argmin(nodes[left, 1])], right[np.argmin(nodes[right, 1])],
                      right[np.argmax(nodes[right, 1])], left[np.argmax(nodes[left, 1])]])
    return Mesh.from_arrays(triangles.nodes, np.asarray(quads), surface_node_ids=triangles.surface_node_ids)


def exact_node_indices(dof_nodes, points, tol: float = 1e-6) -> np.ndarray:
    """Index of the DOF node coinciding with every point (``discretization.py:578-592``)."""

    dof_nodes, points = np.asarray(dof_nodes, dtype=float), np.asarray(points, dtype=float)
    out = np.empty(points.shape[0], dtype=np.int64)
    for i, p in enumerate(points):
        d = np.linalg.norm(dof_nodes - p, axis=1)
        j = int(np.argmin(d))
        if d[j] > tol:
            raise ValueError("field DOF node not found among the primary-potential nodes")
        out[i] = j
    return out


def build(name: str, mesh: Mesh, electrode_positions, wavenumbers, *, refine: bool, quadratic: bool,
          quadrature_order: int = 2) -> dict:
    """One auxiliary discretization (``discretization.py:476-575`` with ``auxiliary=True``) as the
    arrays the operator's device state is built from."""

    strips = structured_quads(mesh)
    parents = np.arange(mesh.cell_count)
    if strips is not None:
        geometry = strips
        if refine:
            geometry, parents = refine_quads(geometry)
        if quadratic:
            dof_nodes, cell_dofs, boundary_dofs = serendipity_topology(geometry)
        else:
            dof_nodes = np.asarray(geometry.nodes, dtype=float)
            cell_dofs = np.asarray(geometry.cells, dtype=np.int64)
            boundary_dofs = np.asarray(geometry.boundary_edges, dtype=np.int64)
        stiffness, mass, boundary_mass = quad_templates(geometry, quadratic)
        electrodes = surface_interpolation(geometry, boundary_dofs if quadratic else None, dof_nodes.shape[0],
                                           electrode_positions)
    else:
        geometry = mesh
        if refine:
            geometry, refined = refine_triangles(geometry)
            parents = parents[refined]
        if quadratic:
            dof_nodes, cell_dofs, boundary_dofs = p2_topology(geometry)
        else:
            dof_nodes = np.asarray(geometry.nodes, dtype=float)
            cell_dofs = np.asarray(geometry.cells, dtype=np.int64)
            boundary_dofs = np.asarray(geometry.boundary_edges, dtype=np.int64)
        stiffness, mass, boundary_mass = triangle_templates(geometry, quadratic, quadrature_order)
        electrodes = point_interpolation(geometry, cell_dofs, dof_nodes.shape[0], electrode_positions)
    center = np.mean(np.asarray(electrode_positions, dtype=float), axis=0)
    alpha = np.stack([fem.robin_2p5d(geometry.boundary_edge_centers, geometry.boundary_edge_normals,
                                     geometry.surface_edge_mask, float(geometry.surface_reference_level), center,
                                     float(k)) for k in np.asarray(wavenumbers, dtype=float)])
    return dict(name=name, geometry=geometry, parents=np.asarray(parents, dtype=np.int64),
                dof_nodes=np.asarray(dof_nodes, dtype=float), cell_dofs=cell_dofs, boundary_dofs=boundary_dofs,
                stiffness=stiffness, mass=mass, boundary_mass=boundary_mass, alpha=alpha, electrodes=electrodes)
