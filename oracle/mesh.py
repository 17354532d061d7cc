"""2D and 3D mesh metadata: boundary topology, surface, point location.

Follows ``adtlert/mesh/core.py:27-247, 543-744`` and ``adtlert/mesh/core3d.py:16-234``.
Test infrastructure only.  Plain dictionaries of float64/int64 arrays.
"""

from __future__ import annotations

import heapq

import numpy as np

TETRA_FACES = ((1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1))  # core3d.py:16
TET_EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))  # fem/tetrahedron.py:13


def _cell_edges(cells):
    """Sorted node pairs of every polygon edge, ``(C, width, 2)`` (``core.py:73-78``)."""

    return np.sort(np.stack((cells, np.roll(cells, -1, axis=1)), axis=-1), axis=-1)


def _surface_path(nodes, edges, tol=1e-8):
    """Dijkstra top-boundary path between the upper corners (``core.py:124-174``)."""

    boundary_nodes = np.unique(edges)
    x = nodes[boundary_nodes, 0]

    def upper_corner(extreme):
        candidates = boundary_nodes[np.abs(x - extreme) <= tol].tolist()
        return int(max(candidates, key=lambda node: (nodes[node, 1], -nodes[node, 0])))

    start, stop = upper_corner(float(x.min())), upper_corner(float(x.max()))
    centers_y = np.mean(nodes[edges], axis=1)[:, 1]
    lengths = np.linalg.norm(nodes[edges[:, 1]] - nodes[edges[:, 0]], axis=1)
    span = max(float(centers_y.max() - centers_y.min()), tol)
    adjacency = {int(n): [] for n in boundary_nodes}
    top = centers_y.max()
    for (a, b), cy, length in zip(edges.tolist(), centers_y, lengths):
        cost = float(length * (1.0 + 100.0 * (top - cy) / span))
        adjacency[a].append((b, cost))
        adjacency[b].append((a, cost))
    dist, prev, seen, heap = {start: 0.0}, {}, set(), [(0.0, start)]
    while heap:
        d, node = heapq.heappop(heap)
        if node in seen:
            continue
        seen.add(node)
        if node == stop:
            break
        for nb, cost in adjacency[node]:
            if d + cost < dist.get(nb, float("inf")):
                dist[nb], prev[nb] = d + cost, node
                heapq.heappush(heap, (d + cost, nb))
    if start == stop:
        return [start]
    if stop not in prev:
        return [start, stop]
    path = [stop]
    while path[-1] != start:
        path.append(prev[path[-1]])
    return path[::-1]


def mesh2d(nodes, cells, surface_node_ids=None, tol=1e-8):
    """``Mesh.from_arrays`` metadata (``core.py:575-630``) as a dict."""

    nodes = np.asarray(nodes, dtype=float)
    cells = np.asarray(cells, dtype=np.int64)
    surface = None if surface_node_ids is None else np.asarray(surface_node_ids, dtype=np.int64)
    used = np.unique(cells)
    if used.size != nodes.shape[0]:
        remap = np.full(nodes.shape[0], -1, dtype=np.int64)
        remap[used] = np.arange(used.size)
        nodes, cells = nodes[used], remap[cells]
        surface = None if surface is None else remap[surface]
    x, y = nodes[cells][..., 0], nodes[cells][..., 1]
    areas = 0.5 * np.abs(np.sum(x * np.roll(y, -1, axis=1) - np.roll(x, -1, axis=1) * y, axis=1))
    edges_all = _cell_edges(cells).reshape(-1, 2)
    unique, first, counts = np.unique(edges_all, axis=0, return_index=True, return_counts=True)
    boundary = counts == 1
    edges, edge_cells = unique[boundary], first[boundary] // cells.shape[1]
    en = nodes[edges]
    centers = en.mean(axis=1)
    vectors = en[:, 1] - en[:, 0]
    lengths = np.linalg.norm(vectors, axis=1)
    normals = np.stack((vectors[:, 1], -vectors[:, 0]), axis=1) / lengths[:, None]
    outward = np.sum(normals * (centers - nodes[cells[edge_cells]].mean(axis=1)), axis=1) >= 0.0
    normals = normals * np.where(outward, 1.0, -1.0)[:, None]
    path = _surface_path(nodes, edges) if surface is None else surface
    path = np.asarray(path, dtype=np.int64)
    surface_pairs = {tuple(sorted(p)) for p in zip(path[:-1].tolist(), path[1:].tolist())}
    mask = np.asarray([tuple(e) in surface_pairs for e in edges.tolist()], dtype=bool)
    snodes = nodes[path]
    flat = bool(np.max(np.abs(snodes[:, 1] - snodes[0, 1])) <= tol) if snodes.size else True
    level = float(np.mean(snodes[:, 1])) if snodes.size else 0.0
    return dict(
        nodes=nodes, cells=cells, areas=areas, boundary_edges=edges, boundary_edge_cells=edge_cells,
        boundary_edge_centers=centers, boundary_edge_lengths=lengths, boundary_edge_normals=normals,
        surface_node_ids=path, surface_edge_mask=mask, surface_level=level, flat=flat,
    )


def locate2d(mesh, points, tol=1e-4):
    """First-hit point location in triangles or Q1 quads (``core.py:199-323``).

    A node within ``tol`` wins (weight 1 in the first cell using it).  Only the
    node-hit and triangle branches are restated; quads are located by node hits
    (every electrode of the quad fixtures sits on a node).
    """

    nodes, cells = mesh["nodes"], mesh["cells"]
    points = np.asarray(points, dtype=float)
    cell_ids = np.full(points.shape[0], -1, dtype=np.int64)
    weights = np.zeros((points.shape[0], cells.shape[1]))
    for p, point in enumerate(points):
        d = np.linalg.norm(nodes - point, axis=1)
        nearest = int(np.argmin(d))
        if d[nearest] <= tol:
            adjacent = np.flatnonzero(np.any(cells == nearest, axis=1))
            cell_ids[p] = adjacent[0]
            weights[p, int(np.flatnonzero(cells[adjacent[0]] == nearest)[0])] = 1.0
            continue
        if cells.shape[1] == 3:
            origins = nodes[cells[:, 0]]
            tr = np.stack((nodes[cells[:, 1]] - origins, nodes[cells[:, 2]] - origins), axis=-1)
            local = np.linalg.solve(tr, (point - origins)[..., None])[..., 0]
            bary = np.column_stack((1.0 - local.sum(axis=1), local))
            hits = np.flatnonzero(np.all((bary >= -tol) & (bary <= 1.0 + tol), axis=1))
            if hits.size:
                cell_ids[p], weights[p] = hits[0], bary[hits[0]]
    if np.any(cell_ids < 0):
        raise ValueError("some points could not be located in the mesh")
    return cell_ids, weights


def midpoint_topology(nodes, cells, faces, cell_edges, face_edges):
    """Shared-edge quadratic DOFs; midpoints numbered in order of first use (``core.py:27-44``)."""

    points = [np.asarray(p, dtype=float) for p in np.asarray(nodes, dtype=float)]
    index = {}

    def midpoint(a, b):
        key = (a, b) if a < b else (b, a)
        if key not in index:
            index[key] = len(points)
            points.append(0.5 * (points[key[0]] + points[key[1]]))
        return index[key]

    quad_cells = [list(c) + [midpoint(c[i], c[j]) for i, j in cell_edges] for c in np.asarray(cells).tolist()]
    quad_faces = [list(f) + [midpoint(f[i], f[j]) for i, j in face_edges] for f in np.asarray(faces).tolist()]
    return np.asarray(points), np.asarray(quad_cells, dtype=np.int64), np.asarray(quad_faces, dtype=np.int64)


def mesh3d(nodes, cells, surface_face_mask=None):
    """``Mesh3D.from_arrays`` metadata (``core3d.py:126-180``) as a dict."""

    nodes = np.asarray(nodes, dtype=float)
    cells = np.asarray(cells, dtype=np.int64)
    used = np.unique(cells)
    if used.size != nodes.shape[0]:
        remap = np.full(nodes.shape[0], -1, dtype=np.int64)
        remap[used] = np.arange(used.size)
        nodes, cells = nodes[used], remap[cells]
    cn = nodes[cells]
    jac = np.stack([cn[:, i] - cn[:, 0] for i in (1, 2, 3)], axis=-1)
    volumes = np.abs(np.linalg.det(jac)) / 6.0
    faces_all = np.sort(cells[:, TETRA_FACES], axis=-1).reshape(-1, 3)
    unique, first, counts = np.unique(faces_all, axis=0, return_index=True, return_counts=True)
    b = counts == 1
    faces, face_cells = unique[b], first[b] // 4
    fn = nodes[faces]
    centers = fn.mean(axis=1)
    cross = np.cross(fn[:, 1] - fn[:, 0], fn[:, 2] - fn[:, 0])
    norm = np.linalg.norm(cross, axis=1)
    normals = cross / norm[:, None]
    outward = np.sum(normals * (centers - nodes[cells[face_cells]].mean(axis=1)), axis=1) >= 0.0
    normals = normals * np.where(outward, 1.0, -1.0)[:, None]
    mask = normals[:, 2] > 1e-8 if surface_face_mask is None else np.asarray(surface_face_mask, bool)
    snodes = np.unique(faces[mask])
    stol = max(1e-8, max(float(np.ptp(nodes[:, 2])), 1.0) * 1e-8)
    return dict(
        nodes=nodes, cells=cells, volumes=volumes, boundary_faces=faces, boundary_face_cells=face_cells,
        boundary_face_centers=centers, boundary_face_areas=0.5 * norm, boundary_face_normals=normals,
        surface_face_mask=mask, surface_level=float(np.max(nodes[:, 2])),
        flat=bool(np.ptp(nodes[snodes, 2]) <= stol),
    )


def locate3d(mesh, points, tol=1e-8):
    """Tetrahedron point location with barycentric weights (``core3d.py:68-106``)."""

    nodes, cells = mesh["nodes"], mesh["cells"]
    points = np.asarray(points, dtype=float)
    cn = nodes[cells]
    origins = cn[:, 0]
    inverse = np.linalg.inv(np.stack([cn[:, i] - cn[:, 0] for i in (1, 2, 3)], axis=-1))
    lower, upper = cn.min(axis=1) - tol, cn.max(axis=1) + tol
    cell_ids = np.full(points.shape[0], -1, dtype=np.int64)
    weights = np.zeros((points.shape[0], 4))
    for p, point in enumerate(points):
        nearest = int(np.argmin(np.linalg.norm(nodes - point, axis=1)))
        if np.linalg.norm(nodes[nearest] - point) <= tol:
            candidates = np.flatnonzero(np.any(cells == nearest, axis=1))
        else:
            candidates = np.flatnonzero(np.all((point >= lower) & (point <= upper), axis=1))
        for c in candidates:
            local = inverse[c] @ (point - origins[c])
            bary = np.asarray((1.0 - np.sum(local), *local))
            if np.all(bary >= -tol) and np.all(bary <= 1.0 + tol):
                bary[np.abs(bary) <= tol] = 0.0
                cell_ids[p], weights[p] = c, bary / np.sum(bary)
                break
    if np.any(cell_ids < 0):
        raise ValueError("points lie outside the tetrahedral mesh")
    return cell_ids, weights
