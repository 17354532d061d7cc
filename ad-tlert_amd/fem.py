"""Unit-conductivity local element templates (vectorized numpy, float64).

The formulas are those of ``adtlert/fem/triangle.py:14-257`` (P1 triangles, 3-point rule),
``adtlert/forward/discretization.py:159-267`` (bilinear Q1 quads, 2x2 Gauss),
``adtlert/fem/boundary.py:12-82`` (2.5D Robin edges with an image source) and
``adtlert/fem/tetrahedron.py:37-127`` plus ``adtlert/forward/ert3d.py:224-256`` (P1/P2
tetrahedra and their far-field Robin faces).  They are exact for the polynomial degrees
involved, so the closed forms pinned in the tests hold to rounding.
"""

from __future__ import annotations

import numpy as np
from scipy.special import k0 as bessel_k0
from scipy.special import k1 as bessel_k1

_P1_POINTS = np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]])
_P1_WEIGHTS = np.full(3, 1 / 6)
_REF_GRAD = np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])


def p1_triangle(nodes: np.ndarray, cells: np.ndarray):
    """P1 stiffness ``|T| grad(l_i).grad(l_j)`` and mass (3-point rule) per cell."""

    cn = nodes[cells]
    jac = np.stack((cn[:, 1] - cn[:, 0], cn[:, 2] - cn[:, 0]), axis=-1)
    grads = np.einsum("eji,nj->eni", np.linalg.inv(jac), _REF_GRAD)
    area = 0.5 * np.abs(np.linalg.det(jac))
    shape = np.stack((1 - _P1_POINTS.sum(axis=1), _P1_POINTS[:, 0], _P1_POINTS[:, 1]), axis=-1)
    qw = 2.0 * area[:, None] * _P1_WEIGHTS[None, :]
    stiffness = np.einsum("eid,ejd,e->eij", grads, grads, qw.sum(axis=1))
    mass = np.einsum("qi,qj,eq->eij", shape, shape, qw)
    return stiffness, mass


def q1_quad(nodes: np.ndarray, cells: np.ndarray):
    """Bilinear quad stiffness and mass with the 2x2 Gauss rule."""

    g = 1.0 / np.sqrt(3.0)
    pts = np.array([[-g, -g], [g, -g], [g, g], [-g, g]])
    xi, eta = pts[:, 0], pts[:, 1]
    values = 0.25 * np.stack(((1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)), -1)
    dxi = 0.25 * np.stack((-(1 - eta), 1 - eta, 1 + eta, -(1 + eta)), axis=-1)
    deta = 0.25 * np.stack((-(1 - xi), -(1 + xi), 1 + xi, 1 - xi), axis=-1)
    grads = np.stack((dxi, deta), axis=-1)
    jac = np.einsum("cid,qia->cqda", nodes[cells], grads)
    phys = np.einsum("qia,cqab->cqib", grads, np.linalg.inv(jac))
    wdet = np.abs(np.linalg.det(jac))
    return np.einsum("cqid,cqjd,cq->cij", phys, phys, wdet), np.einsum("qi,qj,cq->cij", values, values, wdet)


def edge_mass(lengths: np.ndarray) -> np.ndarray:
    return np.asarray(lengths, float)[:, None, None] * (np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0)


def quadratic_edge_mass(lengths: np.ndarray) -> np.ndarray:
    """Robin edge mass of quadratic edges with DOFs (end, midpoint, end) (``fem/boundary.py:12-40``)."""

    return np.asarray(lengths, float)[:, None, None] * (np.array([[4.0, 2.0, -1.0], [2.0, 16.0, 2.0],
                                                                  [-1.0, 2.0, 4.0]]) / 30.0)


def _bilinear(points: np.ndarray):
    xi, eta = points[:, 0], points[:, 1]
    values = 0.25 * np.stack(((1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)), -1)
    dxi = 0.25 * np.stack((-(1 - eta), 1 - eta, 1 + eta, -(1 + eta)), axis=-1)
    deta = 0.25 * np.stack((-(1 - xi), -(1 + xi), 1 + xi, 1 - xi), axis=-1)
    return values, np.stack((dxi, deta), axis=-1)


def serendipity_shape(points: np.ndarray):
    """Eight-node serendipity values and reference gradients on [-1, 1]^2: corners (-1,-1), (1,-1),
    (1,1), (-1,1), then the midpoints of the edges between them (``discretization.py:177-220``)."""

    xi, eta = points[:, 0], points[:, 1]
    values = np.stack((-0.25 * (1 - xi) * (1 - eta) * (1 + xi + eta), -0.25 * (1 + xi) * (1 - eta) * (1 - xi + eta),
                       -0.25 * (1 + xi) * (1 + eta) * (1 - xi - eta), -0.25 * (1 - xi) * (1 + eta) * (1 + xi - eta),
                       0.5 * (1 - xi * xi) * (1 - eta), 0.5 * (1 + xi) * (1 - eta * eta),
                       0.5 * (1 - xi * xi) * (1 + eta), 0.5 * (1 - xi) * (1 - eta * eta)), axis=-1)
    dxi = np.stack((0.25 * (1 - eta) * (2 * xi + eta), 0.25 * (1 - eta) * (2 * xi - eta),
                    0.25 * (1 + eta) * (2 * xi + eta), 0.25 * (1 + eta) * (2 * xi - eta), -xi * (1 - eta),
                    0.5 * (1 - eta * eta), -xi * (1 + eta), -0.5 * (1 - eta * eta)), axis=-1)
    deta = np.stack((0.25 * (1 - xi) * (xi + 2 * eta), 0.25 * (1 + xi) * (-xi + 2 * eta),
                     0.25 * (1 + xi) * (xi + 2 * eta), 0.25 * (1 - xi) * (-xi + 2 * eta), -0.5 * (1 - xi * xi),
                     -(1 + xi) * eta, 0.5 * (1 - xi * xi), -(1 - xi) * eta), axis=-1)
    return values, np.stack((dxi, deta), axis=-1)


def serendipity_quad(nodes: np.ndarray, cells: np.ndarray):
    """Serendipity stiffness and mass on bilinear-geometry quads (corner nodes ``cells[:, :4]``),
    3 x 3 Gauss rule in FP64 (``discretization.py:223-267``)."""

    g, gw = np.polynomial.legendre.leggauss(3)
    pts = np.array([[x, y] for y in g for x in g])
    wts = np.array([wx * wy for wy in gw for wx in gw])
    _, geo = _bilinear(pts)
    values, grads = serendipity_shape(pts)
    jac = np.einsum("cid,qia->cqda", nodes[cells[:, :4]], geo)
    phys = np.einsum("qia,cqab->cqib", grads, np.linalg.inv(jac))
    wdet = np.abs(np.linalg.det(jac)) * wts
    return np.einsum("cqid,cqjd,cq->cij", phys, phys, wdet), np.einsum("qi,qj,cq->cij", values, values, wdet)


# degree-4 six-point rule on the unit triangle (Strang and Fix), weights summing to 1/2
_DEG4_POINTS = np.array([[0.445948490915965, 0.445948490915965], [0.445948490915965, 0.108103018168070],
                         [0.108103018168070, 0.445948490915965], [0.091576213509771, 0.091576213509771],
                         [0.091576213509771, 0.816847572980459], [0.816847572980459, 0.091576213509771]])
_DEG4_WEIGHTS = np.array([0.1116907948390055] * 3 + [0.0549758718276610] * 3)


def p2_triangle(nodes: np.ndarray, cells: np.ndarray):
    """P2 stiffness and mass (vertices, then midpoints 01, 12, 20) with the degree-4 rule
    (``fem/triangle.py:136-257``)."""

    cn = nodes[cells[:, :3]]
    jac = np.stack((cn[:, 1] - cn[:, 0], cn[:, 2] - cn[:, 0]), axis=-1)
    l2, l3 = _DEG4_POINTS[:, 0, None], _DEG4_POINTS[:, 1, None]
    l1 = 1.0 - l2 - l3
    g1, g2, g3 = _REF_GRAD
    ref = np.stack(((4 * l1 - 1) * g1, (4 * l2 - 1) * g2, (4 * l3 - 1) * g3, 4 * (l1 * g2 + l2 * g1),
                    4 * (l2 * g3 + l3 * g2), 4 * (l3 * g1 + l1 * g3)), axis=-2)
    grads = np.einsum("eij,qnj->eqni", np.linalg.inv(jac).transpose(0, 2, 1), ref)
    qw = np.abs(np.linalg.det(jac))[:, None] * _DEG4_WEIGHTS[None, :]
    l1, l2, l3 = l1[:, 0], l2[:, 0], l3[:, 0]
    shape = np.stack((l1 * (2 * l1 - 1), l2 * (2 * l2 - 1), l3 * (2 * l3 - 1), 4 * l1 * l2, 4 * l2 * l3, 4 * l3 * l1), -1)
    return np.einsum("eqid,eqjd,eq->eij", grads, grads, qw), np.einsum("qi,qj,eq->eij", shape, shape, qw)


def robin_2p5d(centers, normals, surface_mask, surface_level, source_center, wavenumber) -> np.ndarray:
    """Unit-conductivity 2.5D mixed-boundary coefficient ``k (r.n/r) K1(kr)/K0(kr)`` with image."""

    centers, normals = np.asarray(centers, float), np.asarray(normals, float)
    s = np.asarray(source_center, float)
    image = np.array([s[0], 2.0 * surface_level - s[1]])
    r1, r2 = centers - s, centers - image
    a1, a2 = np.linalg.norm(r1, axis=1), np.linalg.norm(r2, axis=1)
    out = np.zeros(centers.shape[0])
    valid = np.flatnonzero((a1 > 1e-12) & (a2 > 1e-12) & ~np.asarray(surface_mask, bool))
    denom = bessel_k0(a1[valid] * wavenumber) + bessel_k0(a2[valid] * wavenumber)
    ok = denom > 1e-12
    st = valid[ok]
    num = (np.einsum("bd,bd->b", r1[st], normals[st]) / a1[st] * bessel_k1(a1[st] * wavenumber)
           + np.einsum("bd,bd->b", r2[st], normals[st]) / a2[st] * bessel_k1(a2[st] * wavenumber))
    out[st] = wavenumber * num / denom[ok]
    return out


def tet_gradients(nodes: np.ndarray, cells: np.ndarray):
    cn = nodes[cells[:, :4]]
    jac = np.stack([cn[:, i] - cn[:, 0] for i in (1, 2, 3)], axis=-1)
    ref = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    return np.einsum("ij,ejk->eik", ref, np.linalg.inv(jac)), np.abs(np.linalg.det(jac)) / 6.0


def tet_p1(nodes: np.ndarray, cells: np.ndarray) -> np.ndarray:
    g, vol = tet_gradients(nodes, cells)
    return vol[:, None, None] * np.einsum("eid,ejd->eij", g, g)


def tet_p2(nodes: np.ndarray, cells: np.ndarray) -> np.ndarray:
    """Ten-node stiffness with the 4-point rule (gradients are linear -> exact)."""

    g1, vol = tet_gradients(nodes, cells)
    a, b = 0.585410196624969, 0.138196601125011
    quad = np.array([[a, b, b, b], [b, a, b, b], [b, b, a, b], [b, b, b, a]])
    vertex = (4.0 * quad[None, :, :, None] - 1.0) * g1[:, None, :, :]
    edge = np.stack([4.0 * (quad[None, :, i, None] * g1[:, None, j, :] + quad[None, :, j, None] * g1[:, None, i, :])
                     for i, j in ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))], axis=2)
    g = np.concatenate((vertex, edge), axis=2)
    return vol[:, None, None] * np.einsum("eqid,eqjd->eij", g, g) / 4.0


def tet_p2_shape(bary: np.ndarray) -> np.ndarray:
    bary = np.asarray(bary, float)
    edges = np.stack([4.0 * bary[..., i] * bary[..., j] for i, j in ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))], -1)
    return np.concatenate((bary * (2.0 * bary - 1.0), edges), axis=-1)


def p2_face_mass_unit() -> np.ndarray:
    """Six-node triangle mass per unit area by the degree-4 rule (``tetrahedron.py:69-90``)."""

    a, b, c, d = 0.445948490915965, 0.108103018168070, 0.091576213509771, 0.816847572980459
    bary = np.array([(a, a, b), (a, b, a), (b, a, a), (c, c, d), (c, d, c), (d, c, c)])
    w = np.array((0.223381589678011,) * 3 + (0.109951743655322,) * 3)
    edges = np.stack([4.0 * bary[:, i] * bary[:, j] for i, j in ((0, 1), (1, 2), (2, 0))], -1)
    shape = np.concatenate((bary * (2.0 * bary - 1.0), edges), axis=1)
    return np.einsum("q,qi,qj->ij", w, shape, shape)


P1_FACE_MASS_UNIT = np.array([[2.0, 1.0, 1.0], [1.0, 2.0, 1.0], [1.0, 1.0, 2.0]]) / 12.0


def add_robin_faces(cell_templates, conn, bconn, face_cells, centers, areas, normals, surface_mask, source_center,
                    order: int) -> None:
    """Add ``alpha * face_mass`` of every non-surface boundary face into its cell's template,
    ``alpha = max(r.n / |r|^2, 0)`` with ``r`` from the electrode centroid (``ert3d.py:224-256``)."""

    far = np.flatnonzero(~np.asarray(surface_mask, bool))
    radial = np.asarray(centers, float)[far] - np.asarray(source_center, float)
    r2 = np.einsum("fd,fd->f", radial, radial)
    alpha = np.zeros_like(r2)
    ok = r2 > 0.0
    alpha[ok] = np.maximum(np.einsum("fd,fd->f", radial[ok], np.asarray(normals, float)[far][ok]) / r2[ok], 0.0)
    keep = alpha != 0.0
    far, alpha = far[keep], alpha[keep]
    cells = np.asarray(face_cells)[far]
    fdofs = np.asarray(bconn)[far]
    cdofs = np.asarray(conn)[cells]
    local = np.argmax(cdofs[:, None, :] == fdofs[:, :, None], axis=2)  # position of each face dof in its cell
    unit = p2_face_mass_unit() if order == 2 else P1_FACE_MASS_UNIT
    blocks = (alpha * np.asarray(areas, float)[far])[:, None, None] * unit[None]
    kf = fdofs.shape[1]
    ci = np.repeat(cells, kf * kf)
    li = np.repeat(local, kf, axis=1).reshape(-1)
    lj = np.tile(local, (1, kf)).reshape(-1)
    np.add.at(cell_templates, (ci, li, lj), blocks.reshape(-1))
