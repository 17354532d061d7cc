"""Local element templates and boundary terms (float64).

Follows ``adtlert/fem/triangle.py:14-257``, ``adtlert/fem/tetrahedron.py:13-127``,
``adtlert/fem/boundary.py:12-82``, ``adtlert/forward/discretization.py:159-285``
and the 3D Robin faces of ``adtlert/forward/ert3d.py:224-256``.  Test infrastructure only.
"""

from __future__ import annotations

import numpy as np
from scipy.special import k0 as bessel_k0
from scipy.special import k1 as bessel_k1

P1_RULE = (np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]]), np.array([1 / 6] * 3))  # triangle.py:14-20
DEG4_POINTS = np.array([
    [0.445948490915965, 0.445948490915965], [0.445948490915965, 0.108103018168070],
    [0.108103018168070, 0.445948490915965], [0.091576213509771, 0.091576213509771],
    [0.091576213509771, 0.816847572980459], [0.816847572980459, 0.091576213509771],
])  # triangle.py:21-31
DEG4_WEIGHTS = np.array([0.1116907948390055] * 3 + [0.0549758718276610] * 3)
REF_GRAD = np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])


def _tri_jacobians(nodes, cells):
    cn = nodes[cells[:, :3]]
    return np.stack((cn[:, 1] - cn[:, 0], cn[:, 2] - cn[:, 0]), axis=-1)


def p1_triangle(nodes, cells):
    """P1 stiffness and mass, 3-point rule (``triangle.py:94-133, 203-235``)."""

    jac = _tri_jacobians(nodes, cells)
    grads = np.einsum("eij,nj->eni", np.linalg.inv(jac).transpose(0, 2, 1), REF_GRAD)
    areas = 0.5 * np.abs(np.linalg.det(jac))
    pts, wts = P1_RULE
    shape = np.stack((1 - pts[:, 0] - pts[:, 1], pts[:, 0], pts[:, 1]), axis=-1)
    qw = 2.0 * areas[:, None] * wts[None, :]
    stiffness = np.einsum("eid,ejd,eq->eij", grads, grads, qw)
    mass = np.einsum("qi,qj,eq->eij", shape, shape, qw)
    return stiffness, mass


def p2_triangle_shape(points):
    l2, l3 = points[:, 0], points[:, 1]
    l1 = 1.0 - l2 - l3
    return np.stack((l1 * (2 * l1 - 1), l2 * (2 * l2 - 1), l3 * (2 * l3 - 1), 4 * l1 * l2, 4 * l2 * l3, 4 * l3 * l1), -1)


def p2_triangle(nodes, cells):
    """P2 stiffness and mass, degree-4 rule (``triangle.py:136-257``)."""

    jac = _tri_jacobians(nodes, cells)
    l2, l3 = DEG4_POINTS[:, 0, None], DEG4_POINTS[:, 1, None]
    l1 = 1.0 - l2 - l3
    g1, g2, g3 = REF_GRAD
    ref = np.stack(((4 * l1 - 1) * g1, (4 * l2 - 1) * g2, (4 * l3 - 1) * g3,
                    4 * (l1 * g2 + l2 * g1), 4 * (l2 * g3 + l3 * g2), 4 * (l3 * g1 + l1 * g3)), axis=-2)
    grads = np.einsum("eij,qnj->eqni", np.linalg.inv(jac).transpose(0, 2, 1), ref)
    areas = 0.5 * np.abs(np.linalg.det(jac))
    qw = 2.0 * areas[:, None] * DEG4_WEIGHTS[None, :]
    shape = p2_triangle_shape(DEG4_POINTS)
    return np.einsum("eqid,eqjd,eq->eij", grads, grads, qw), np.einsum("qi,qj,eq->eij", shape, shape, qw)


def bilinear(points):
    """Q1 values and reference gradients on [-1, 1]^2 (``discretization.py:159-174``)."""

    xi, eta = points[:, 0], points[:, 1]
    values = 0.25 * np.stack(((1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)), -1)
    dxi = 0.25 * np.stack((-(1 - eta), 1 - eta, 1 + eta, -(1 + eta)), axis=-1)
    deta = 0.25 * np.stack((-(1 - xi), -(1 + xi), 1 + xi, 1 - xi), axis=-1)
    return values, np.stack((dxi, deta), axis=-1)


def q1_quad(nodes, cells):
    """Bilinear quad stiffness and mass, 2x2 Gauss (``discretization.py:223-267``)."""

    g = 1.0 / np.sqrt(3.0)
    points = np.array([[-g, -g], [g, -g], [g, g], [-g, g]])
    values, grads = bilinear(points)
    cn = nodes[cells]
    jac = np.einsum("cid,qia->cqda", cn, grads)
    phys = np.einsum("qia,cqab->cqib", grads, np.linalg.inv(jac))
    wdet = np.abs(np.linalg.det(jac))
    return np.einsum("cqid,cqjd,cq->cij", phys, phys, wdet), np.einsum("qi,qj,cq->cij", values, values, wdet)


def edge_mass(lengths):
    """Linear Robin edge matrices ``L/6 [[2,1],[1,2]]`` (``boundary.py:31-34``)."""

    return np.asarray(lengths, dtype=float)[:, None, None] * (np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0)


def robin_2p5d(mesh, center, wavenumber):
    """2.5D mixed-boundary coefficients with an image source (``boundary.py:43-82``), unit conductivity."""

    centers, normals = mesh["boundary_edge_centers"], mesh["boundary_edge_normals"]
    source = np.asarray(center, dtype=float)
    image = np.array([source[0], 2.0 * mesh["surface_level"] - source[1]])
    r1, r2 = centers - source, centers - image
    a1, a2 = np.linalg.norm(r1, axis=1), np.linalg.norm(r2, axis=1)
    out = np.zeros(centers.shape[0])
    valid = np.flatnonzero((a1 > 1e-12) & (a2 > 1e-12) & ~mesh["surface_edge_mask"])
    denom = bessel_k0(a1[valid] * wavenumber) + bessel_k0(a2[valid] * wavenumber)
    ok = denom > 1e-12
    stable = valid[ok]
    num = (np.sum(r1[stable] * normals[stable], axis=1) / a1[stable] * bessel_k1(a1[stable] * wavenumber)
           + np.sum(r2[stable] * normals[stable], axis=1) / a2[stable] * bessel_k1(a2[stable] * wavenumber))
    out[stable] = wavenumber * num / denom[ok]
    return out


def tet_p1(nodes, cells):
    """Constant P1 gradients, volumes and stiffness (``tetrahedron.py:37-56``)."""

    cn = nodes[cells[:, :4]]
    jac = np.stack([cn[:, i] - cn[:, 0] for i in (1, 2, 3)], axis=-1)
    ref = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    grads = np.einsum("ij,ejk->eik", ref, np.linalg.inv(jac))
    volumes = np.abs(np.linalg.det(jac)) / 6.0
    return grads, volumes, volumes[:, None, None] * np.einsum("eid,ejd->eij", grads, grads)


def tet_p2_shape(bary):
    """Ten-node shape values, vertices then edges 01,12,20,03,13,23 (``tetrahedron.py:59-66``)."""

    bary = np.asarray(bary, dtype=float)
    edges = np.stack([4.0 * bary[..., i] * bary[..., j] for i, j in ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))], -1)
    return np.concatenate((bary * (2.0 * bary - 1.0), edges), axis=-1)


def tet_p2_stiffness(nodes, cells):
    """Ten-node stiffness, 4-point rule (``tetrahedron.py:93-119``)."""

    grads1, volumes, _ = tet_p1(nodes, cells)
    a, b = 0.585410196624969, 0.138196601125011
    quad = np.array([[a, b, b, b], [b, a, b, b], [b, b, a, b], [b, b, b, a]])
    vertex = (4.0 * quad[None, :, :, None] - 1.0) * grads1[:, None, :, :]
    edge = np.stack([4.0 * (quad[None, :, i, None] * grads1[:, None, j, :] + quad[None, :, j, None] * grads1[:, None, i, :])
                     for i, j in ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))], axis=2)
    g = np.concatenate((vertex, edge), axis=2)
    return volumes[:, None, None] * np.einsum("eqid,eqjd->eij", g, g) / 4.0


def p2_face_mass(area):
    """Six-node triangle mass by the degree-4 rule (``tetrahedron.py:69-90``)."""

    a, b, c, d = 0.445948490915965, 0.108103018168070, 0.091576213509771, 0.816847572980459
    bary = np.array([(a, a, b), (a, b, a), (b, a, a), (c, c, d), (c, d, c), (d, c, c)])
    w = np.array((0.223381589678011,) * 3 + (0.109951743655322,) * 3)
    edges = np.stack([4.0 * bary[:, i] * bary[:, j] for i, j in ((0, 1), (1, 2), (2, 0))], -1)
    shape = np.concatenate((bary * (2.0 * bary - 1.0), edges), axis=1)
    return float(area) * np.einsum("q,qi,qj->ij", w, shape, shape)


def robin_coefficient_3d(centers, normals, center):
    """First-order far-field coefficient ``alpha = max(r . n / |r|^2, 0)`` with ``r`` from the
    survey centre to each face centre (``ert3d.py:239-256``); faces at the centre get 0."""

    radial = np.asarray(centers, float) - np.asarray(center, float)
    r2 = np.einsum("fd,fd->f", radial, radial)
    dot = np.einsum("fd,fd->f", radial, np.asarray(normals, float))
    out = np.zeros(radial.shape[0])
    ok = r2 > 0.0
    out[ok] = np.maximum(dot[ok] / r2[ok], 0.0)
    return out


P1_FACE_MASS = np.array([[2.0, 1.0, 1.0], [1.0, 2.0, 1.0], [1.0, 1.0, 2.0]]) / 12.0  # ert3d.py:236-238
