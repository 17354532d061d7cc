"""2D triangle/quadrilateral and 3D tetrahedral meshes with boundary and surface metadata.

Interface of ``adtlert/mesh/core.py:557-744`` (``Mesh``) and ``adtlert/mesh/core3d.py:109-234``
(``Mesh3D``) for the flat-surface forward path: ``from_arrays``, boundary edges/faces with
owning cells, centers, lengths/areas and outward normals, the ground surface, point location
and quadratic topology.  Topology is built with vectorized integer keys (one ``np.unique`` over
int64 codes) so million-cell meshes preprocess in seconds.
"""

from __future__ import annotations

import heapq
from dataclasses import dataclass

import numpy as np
import torch

from .dtypes import FLOAT_DTYPE, INT_DTYPE, NP_FLOAT_DTYPE

Tensor = torch.Tensor
TETRA_FACES = ((1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1))
TET_EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))


def _float(a) -> Tensor:
    return torch.as_tensor(np.asarray(a, dtype=NP_FLOAT_DTYPE), dtype=FLOAT_DTYPE)


def _int(a) -> Tensor:
    return torch.as_tensor(np.asarray(a, dtype=np.int32).copy(), dtype=INT_DTYPE)


def _drop_unused(nodes, cells, extra=None):
    used = np.unique(cells)
    if used.size == nodes.shape[0]:
        return nodes, cells, extra
    remap = np.full(nodes.shape[0], -1, dtype=np.int64)
    remap[used] = np.arange(used.size)
    return nodes[used], remap[cells], None if extra is None else remap[extra]


def _surface_path(nodes: np.ndarray, edges: np.ndarray, tol: float = 1e-8) -> list[int]:
    """Top boundary path between the upper-left and upper-right corners (Dijkstra with costs
    growing with depth below the highest boundary edge), as ``mesh/core.py:124-174``."""

    boundary_nodes = np.unique(edges)
    x = nodes[boundary_nodes, 0]

    def upper_corner(extreme: float) -> int:
        candidates = boundary_nodes[np.abs(x - extreme) <= tol].tolist()
        return int(max(candidates, key=lambda n: (nodes[n, 1], -nodes[n, 0])))

    start, stop = upper_corner(float(x.min())), upper_corner(float(x.max()))
    cy = nodes[edges].mean(axis=1)[:, 1]
    lengths = np.linalg.norm(nodes[edges[:, 1]] - nodes[edges[:, 0]], axis=1)
    span = max(float(cy.max() - cy.min()), tol)
    adjacency: dict[int, list[tuple[int, float]]] = {int(n): [] for n in boundary_nodes}
    top = float(cy.max())
    for (a, b), y, length in zip(edges.tolist(), cy.tolist(), lengths.tolist()):
        cost = float(length * (1.0 + 100.0 * (top - y) / span))
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


def _bilinear_weights(xi: float, eta: float):
    values = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
    grads = 0.25 * np.stack((np.array([-(1 - eta), 1 - eta, 1 + eta, -(1 + eta)]),
                             np.array([-(1 - xi), -(1 + xi), 1 + xi, 1 - xi])), axis=1)
    return values, grads


def _quad_weights(quad: np.ndarray, point: np.ndarray, tol: float):
    """Q1 weights of a point on a quad edge (linear) or inside it (Newton), ``mesh/core.py:269-306``."""

    for start in range(4):
        stop = (start + 1) % 4
        edge = quad[stop] - quad[start]
        length_sq = float(edge @ edge)
        if length_sq <= tol * tol:
            continue
        t = float((point - quad[start]) @ edge / length_sq)
        if -tol <= t <= 1.0 + tol:
            t = min(max(t, 0.0), 1.0)
            if float(np.linalg.norm(point - (quad[start] + t * edge))) <= tol:
                w = np.zeros(4)
                w[start], w[stop] = 1.0 - t, t
                return w
    xi = eta = 0.0
    for _ in range(32):
        values, grads = _bilinear_weights(xi, eta)
        residual = values @ quad - point
        if float(np.linalg.norm(residual)) <= tol and -1 - tol <= xi <= 1 + tol and -1 - tol <= eta <= 1 + tol:
            break
        try:
            step = np.linalg.solve(grads.T @ quad, residual)
        except np.linalg.LinAlgError:
            return None
        xi, eta = xi - float(step[0]), eta - float(step[1])
        if float(np.linalg.norm(step)) <= 1e-12:
            break
    if not (-1 - tol <= xi <= 1 + tol and -1 - tol <= eta <= 1 + tol):
        return None
    values, _ = _bilinear_weights(min(max(xi, -1.0), 1.0), min(max(eta, -1.0), 1.0))
    return None if np.any(values < -tol) or np.any(values > 1 + tol) else values


@dataclass(frozen=True)
class Mesh:
    """2D triangle or quadrilateral mesh (``mesh/core.py:557-630``)."""

    nodes: Tensor
    cells: Tensor
    boundary_edges: Tensor
    boundary_edge_cells: Tensor
    boundary_edge_centers: Tensor
    boundary_edge_lengths: Tensor
    boundary_edge_normals: Tensor
    surface_node_ids: Tensor
    surface_edge_mask: Tensor
    surface_nodes: Tensor
    surface_reference_level: Tensor
    flat_surface: Tensor
    cell_areas: Tensor

    @classmethod
    def from_arrays(cls, nodes, cells, *, surface_node_ids=None) -> "Mesh":
        nodes = np.asarray(_float(nodes))
        cells = np.asarray(cells, dtype=np.int64)
        surface = None if surface_node_ids is None else np.asarray(surface_node_ids, dtype=np.int64)
        if nodes.ndim != 2 or nodes.shape[1] != 2:
            raise ValueError("nodes must have shape (num_nodes, 2)")
        if cells.ndim != 2 or cells.shape[1] not in (3, 4):
            raise ValueError("cells must have shape (num_cells, 3) or (num_cells, 4)")
        if np.any(cells < 0):
            raise ValueError("cells contain negative node indices")
        if cells.size and np.any(cells >= nodes.shape[0]):
            raise ValueError("cells reference nodes outside the mesh")
        nodes, cells, surface = _drop_unused(nodes, cells, surface)
        xy = nodes[cells].astype(NP_FLOAT_DTYPE)
        x, y = xy[..., 0], xy[..., 1]
        areas = 0.5 * np.abs(np.sum(x * np.roll(y, -1, axis=1) - np.roll(x, -1, axis=1) * y, axis=1))
        if np.any(areas <= 0.0):
            raise ValueError("cells must define non-degenerate 2D polygons")
        n = nodes.shape[0]
        width = cells.shape[1]
        pairs = np.sort(np.stack((cells, np.roll(cells, -1, axis=1)), axis=-1), axis=-1).reshape(-1, 2)
        keys = pairs[:, 0] * n + pairs[:, 1]
        _, first, counts = np.unique(keys, return_index=True, return_counts=True)
        b = first[counts == 1]
        edges, edge_cells = pairs[b], b // width
        en = nodes[edges]
        centers = en.mean(axis=1)
        vec = en[:, 1] - en[:, 0]
        lengths = np.linalg.norm(vec, axis=1)
        normals = np.stack((vec[:, 1], -vec[:, 0]), axis=1) / lengths[:, None]
        outward = np.sum(normals * (centers - nodes[cells[edge_cells]].mean(axis=1)), axis=1) >= 0.0
        normals = normals * np.where(outward, 1.0, -1.0).astype(NP_FLOAT_DTYPE)[:, None]
        nodes64 = nodes.astype(float)
        path = np.asarray(_surface_path(nodes64, edges) if surface is None else surface, dtype=np.int64)
        pkeys = np.sort(np.stack((path[:-1], path[1:]), axis=1), axis=1) if path.size > 1 else np.zeros((0, 2), np.int64)
        mask = np.isin(edges[:, 0] * n + edges[:, 1], pkeys[:, 0] * n + pkeys[:, 1])
        snodes = nodes64[path]
        flat = bool(np.max(np.abs(snodes[:, 1] - snodes[0, 1])) <= 1e-8) if snodes.size else True
        level = float(np.mean(snodes[:, 1])) if snodes.size else 0.0
        return cls(nodes=_float(nodes), cells=_int(cells), boundary_edges=_int(edges), boundary_edge_cells=_int(edge_cells),
                   boundary_edge_centers=_float(centers), boundary_edge_lengths=_float(lengths),
                   boundary_edge_normals=_float(normals), surface_node_ids=_int(path),
                   surface_edge_mask=torch.as_tensor(mask), surface_nodes=_float(snodes),
                   surface_reference_level=torch.tensor(level, dtype=FLOAT_DTYPE), flat_surface=torch.tensor(flat),
                   cell_areas=_float(areas))

    @property
    def node_count(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def cell_count(self) -> int:
        return int(self.cells.shape[0])

    @property
    def cell_node_count(self) -> int:
        return int(self.cells.shape[1])

    @property
    def is_triangle_mesh(self) -> bool:
        return self.cell_node_count == 3

    @property
    def is_quadrilateral_mesh(self) -> bool:
        return self.cell_node_count == 4

    @property
    def is_flat_surface(self) -> bool:
        return bool(self.flat_surface)

    def locate_points(self, points, tol: float = 1e-4) -> tuple[Tensor, Tensor]:
        """Containing cell ids and interpolation weights (node within ``tol`` wins), ``mesh/core.py:199-323``."""

        nodes, cells = np.asarray(self.nodes, dtype=float), np.asarray(self.cells, dtype=np.int64)
        points = np.asarray(points, dtype=float)
        ids = np.full(points.shape[0], -1, dtype=np.int64)
        weights = np.zeros((points.shape[0], cells.shape[1]))
        if self.is_triangle_mesh:
            origins = nodes[cells[:, 0]]
            transforms = np.stack((nodes[cells[:, 1]] - origins, nodes[cells[:, 2]] - origins), axis=-1)
        for p, point in enumerate(points):
            d = np.linalg.norm(nodes - point, axis=1)
            nearest = int(np.argmin(d))
            if d[nearest] <= tol:
                adjacent = np.flatnonzero(np.any(cells == nearest, axis=1))
                if adjacent.size:
                    ids[p] = adjacent[0]
                    weights[p, int(np.flatnonzero(cells[adjacent[0]] == nearest)[0])] = 1.0
                    continue
            if self.is_triangle_mesh:
                local = np.linalg.solve(transforms, (point - origins)[..., None])[..., 0]
                bary = np.column_stack((1.0 - local.sum(axis=1), local))
                hits = np.flatnonzero(np.all((bary >= -tol) & (bary <= 1.0 + tol), axis=1))
                if hits.size:
                    ids[p], weights[p] = hits[0], bary[hits[0]]
            else:
                for c, cell in enumerate(cells):
                    w = _quad_weights(nodes[cell], point, tol)
                    if w is not None:
                        ids[p], weights[p] = c, w
                        break
        if np.any(ids < 0):
            raise ValueError("some points could not be located in the mesh")
        return _int(ids), _float(weights)


def _quadratic_numbering(n_nodes, cells, local_edges):
    """Edge-midpoint dofs numbered in order of first use while walking cells (``core.py:27-44``)."""

    pairs = np.stack([np.sort(cells[:, [i, j]], axis=1) for i, j in local_edges], axis=1).reshape(-1, 2)
    keys = pairs[:, 0] * n_nodes + pairs[:, 1]
    unique, first, inverse = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    mid = (n_nodes + rank[inverse.reshape(-1)]).reshape(cells.shape[0], len(local_edges))
    return unique[order], mid, unique, rank


@dataclass(frozen=True)
class Mesh3D:
    """Four-node tetrahedral mesh with boundary-face metadata (``mesh/core3d.py:109-180``)."""

    nodes: Tensor
    cells: Tensor
    boundary_faces: Tensor
    boundary_face_cells: Tensor
    boundary_face_centers: Tensor
    boundary_face_areas: Tensor
    boundary_face_normals: Tensor
    surface_face_mask: Tensor
    surface_reference_level: Tensor
    flat_surface: Tensor
    cell_volumes: Tensor

    @classmethod
    def from_arrays(cls, nodes, cells, *, surface_face_mask=None) -> "Mesh3D":
        nodes = np.asarray(nodes, dtype=NP_FLOAT_DTYPE)
        cells = np.asarray(cells, dtype=np.int64)
        if nodes.ndim != 2 or nodes.shape[1] != 3:
            raise ValueError("3D nodes must have shape (num_nodes, 3)")
        if cells.ndim != 2 or cells.shape[1] != 4:
            raise ValueError("3D cells must have shape (num_cells, 4)")
        if cells.size == 0:
            raise ValueError("tetrahedral mesh must contain at least one cell")
        if np.any(cells < 0) or np.any(cells >= nodes.shape[0]):
            raise ValueError("cells reference nodes outside the mesh")
        nodes, cells, _ = _drop_unused(nodes, cells)
        cn = nodes[cells]
        jac = np.stack([cn[:, i] - cn[:, 0] for i in (1, 2, 3)], axis=-1)
        volumes = np.abs(np.linalg.det(jac.astype(float))) / 6.0
        if np.any(volumes <= 0.0):
            raise ValueError("cells must define non-degenerate tetrahedra")
        n = nodes.shape[0]
        faces_all = np.sort(cells[:, TETRA_FACES], axis=-1).reshape(-1, 3)
        keys = (faces_all[:, 0] * n + faces_all[:, 1]) * n + faces_all[:, 2]
        _, first, counts = np.unique(keys, return_index=True, return_counts=True)
        b = first[counts == 1]
        faces, face_cells = faces_all[b], b // 4
        fn = nodes[faces].astype(float)
        centers = fn.mean(axis=1)
        cross = np.cross(fn[:, 1] - fn[:, 0], fn[:, 2] - fn[:, 0])
        norm = np.linalg.norm(cross, axis=1)
        normals = cross / norm[:, None]
        outward = np.sum(normals * (centers - nodes[cells[face_cells]].astype(float).mean(axis=1)), axis=1) >= 0.0
        normals = normals * np.where(outward, 1.0, -1.0)[:, None]
        if surface_face_mask is None:
            mask = normals[:, 2] > 1e-8
        else:
            mask = np.asarray(surface_face_mask, dtype=bool).reshape(-1)
            if mask.shape != (faces.shape[0],):
                raise ValueError(f"surface_face_mask must have one entry per boundary face ({mask.shape} != ({faces.shape[0]},))")
        if not np.any(mask):
            raise ValueError("could not identify any upward-facing terrain surface faces")
        snodes = np.unique(faces[mask])
        stol = max(1e-8, max(float(np.ptp(nodes[:, 2])), 1.0) * 1e-8)
        return cls(nodes=_float(nodes), cells=_int(cells), boundary_faces=_int(faces), boundary_face_cells=_int(face_cells),
                   boundary_face_centers=_float(centers), boundary_face_areas=_float(0.5 * norm),
                   boundary_face_normals=_float(normals), surface_face_mask=torch.as_tensor(mask),
                   surface_reference_level=torch.tensor(float(np.max(nodes[:, 2])), dtype=FLOAT_DTYPE),
                   flat_surface=torch.tensor(bool(np.ptp(nodes[snodes, 2]) <= stol)), cell_volumes=_float(volumes))

    @property
    def node_count(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def cell_count(self) -> int:
        return int(self.cells.shape[0])

    @property
    def dimension(self) -> int:
        return 3

    @property
    def is_flat_surface(self) -> bool:
        return bool(self.flat_surface)

    def locate_points(self, points, tol: float = 1e-8) -> tuple[Tensor, Tensor]:
        """Containing tetrahedra and barycentric weights (``mesh/core3d.py:68-106``)."""

        nodes, cells = np.asarray(self.nodes, dtype=float), np.asarray(self.cells, dtype=np.int64)
        points = np.asarray(points, dtype=float)
        cn = nodes[cells]
        lower, upper = cn.min(axis=1) - tol, cn.max(axis=1) + tol
        ids = np.full(points.shape[0], -1, dtype=np.int64)
        weights = np.zeros((points.shape[0], 4))
        for p, point in enumerate(points):
            nearest = int(np.argmin(np.linalg.norm(nodes - point, axis=1)))
            if np.linalg.norm(nodes[nearest] - point) <= tol:
                candidates = np.flatnonzero(np.any(cells == nearest, axis=1))
            else:
                candidates = np.flatnonzero(np.all((point >= lower) & (point <= upper), axis=1))
            for c in candidates:
                jac = np.stack([cn[c, i] - cn[c, 0] for i in (1, 2, 3)], axis=-1)
                local = np.linalg.solve(jac, point - cn[c, 0])
                bary = np.asarray((1.0 - np.sum(local), *local))
                if np.all(bary >= -tol) and np.all(bary <= 1.0 + tol):
                    bary[np.abs(bary) <= tol] = 0.0
                    ids[p], weights[p] = c, bary / np.sum(bary)
                    break
        if np.any(ids < 0):
            raise ValueError(f"points lie outside the tetrahedral mesh: indices={np.flatnonzero(ids < 0).tolist()}")
        return _int(ids), _float(weights)

    def build_quadratic_topology(self) -> tuple[Tensor, Tensor, Tensor]:
        """P2 nodes, cell dofs (vertices, then edges 01,12,20,03,13,23) and boundary-face dofs."""

        nodes = np.asarray(self.nodes, dtype=float)
        cells = np.asarray(self.cells, dtype=np.int64)
        faces = np.asarray(self.boundary_faces, dtype=np.int64)
        n = nodes.shape[0]
        ordered_keys, mid, sorted_keys, rank = _quadratic_numbering(n, cells, TET_EDGES)
        a, b = ordered_keys // n, ordered_keys % n
        points = np.concatenate((nodes, 0.5 * (nodes[a] + nodes[b])))
        fpairs = np.stack([np.sort(faces[:, [i, j]], axis=1) for i, j in ((0, 1), (1, 2), (2, 0))], axis=1)
        fkeys = fpairs[..., 0] * n + fpairs[..., 1]
        fmid = n + rank[np.searchsorted(sorted_keys, fkeys)]
        return _float(points), _int(np.concatenate((cells, mid), axis=1)), _int(np.concatenate((faces, fmid), axis=1))
