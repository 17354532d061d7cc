"""Pinning the FP64 CPU oracle (CPU only).

Each oracle function is checked against something that does not restate its formula: the
reference suite's known-answer tests (geometric factors ``tests/test_survey.py:15-26``,
Wenner ``k = 2 pi a``, mesh counts ``tests/test_mesh_fem.py:37-56``, unit-tet volume), closed
forms (P1/Q1 element matrices, Gauss-rule exactness of the wavenumber quadrature, the Bessel
cosine-transform identity), independent high-order quadrature (P2 templates), finite
differences (Robin coefficients, Jacobians) and the physics invariants the reference tests
(homogeneous half space, 1/sigma scaling, reciprocity, polarity, pole differences, dot tests).
"""

import math

import numpy as np
import pytest
from scipy.special import k0 as K0

from oracle import cpu_pcg, ert2p5d, ert3d, fem, mesh as omesh, quadrature, survey as osurvey
from tests.conftest import adt, relmax

R = -1


# ---------------------------------------------------------------- survey & quadrature


def test_geometric_factors_known_answers():
    line = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [3.0, 0.0]])
    quads = [[0, 1, 2, 3], [0, R, 1, 2], [0, 1, 2, R], [0, R, 1, R]]
    expected = [2 * math.pi / (1 / 2 - 1 / 3 - 1 + 1 / 2), 2 * math.pi / (1 - 1 / 2), 2 * math.pi / (1 / 2 - 1),
                2 * math.pi]
    np.testing.assert_allclose(osurvey.geometric_factors(line, quads), expected, rtol=1e-12)
    assert osurvey.measurement_slots(quads, 4).tolist()[3] == [0, 4, 1, 4]


def test_wenner_alpha_rows_and_factor():
    rows = osurvey.wenner_alpha(10)
    assert len(rows) == sum(10 - 3 * s for s in range(1, 4))
    positions = np.column_stack((np.arange(10.0), np.zeros(10)))
    k = osurvey.geometric_factors(positions, rows)
    spacing = (rows[:, 2] - rows[:, 0]).astype(float)
    np.testing.assert_allclose(k, 2 * math.pi * spacing, rtol=1e-12)


@pytest.mark.parametrize("quad, message", [([R, R, 1, 2], "needs an electrode"), ([0, 1, R, R], "needs an electrode"),
                                           ([0, -2, 1, 2], "other than -1"), ([0, 0, 1, 2], "distinct"),
                                           ([0, 1, 2, 4], "outside the survey")])
def test_invalid_measurements_are_rejected(quad, message):
    line = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [3.0, 0.0]])
    with pytest.raises(ValueError, match=message):
        osurvey.validate(line, [quad])


def test_quadrature_gauss_exactness_pins_points_and_weights():
    r_min, r_max = 1.0, 200.0
    k, q = quadrature.inverse_cosine_weights(r_min, r_max)
    n = max(int(6 * np.log10(r_max / r_min)), 4)
    k0 = 1 / (2 * r_min)
    head, tail = slice(0, n), slice(n, None)
    for j in range(n):  # sqrt(k)-Legendre: exact for k^j, j < n
        assert math.isclose(np.pi * np.sum(q[head] * k[head] ** j), k0 ** (j + 1) / (j + 1), rel_tol=1e-12)
    t = k[tail] / k0 - 1.0
    for j in range(8):  # Laguerre tail: exact for exp(-t) t^j, j <= 7
        assert math.isclose(np.pi * np.sum(q[tail] * np.exp(-t) * t ** j), k0 * math.factorial(j), rel_tol=1e-10)


def test_quadrature_inverts_the_half_space_cosine_transform():
    # (1/pi) int_0^inf K0(k r)/pi dk = 1/(2 pi r): the 2.5D surface potential of a unit source
    k, q = quadrature.inverse_cosine_weights(1.0, 96.0)
    for r in (1.0, 3.0, 10.0, 40.0):
        approx = np.sum(q * K0(k * r) / np.pi)
        assert abs(approx * 2 * np.pi * r - 1.0) < 2e-2


# ---------------------------------------------------------------- meshes


def test_triangle_mesh_topology_known_answers():
    nodes, cells = adt.cases.grid_triangles()
    m = omesh.mesh2d(nodes, cells)
    assert cells.shape[0] == 48 and m["boundary_edges"].shape[0] == 20
    assert math.isclose(m["areas"].sum(), 24.0)
    assert m["flat"] and np.allclose(nodes[m["surface_node_ids"], 1], 0.0)
    centers = nodes[cells].mean(axis=1)[m["boundary_edge_cells"]]
    assert np.all(np.sum(m["boundary_edge_normals"] * (m["boundary_edge_centers"] - centers), axis=1) > 0)
    assert not omesh.mesh2d(*adt.cases.grid_triangles(slope=0.2))["flat"]


def test_locate_points_reproduces_coordinates():
    nodes, cells = adt.cases.grid_triangles()
    m = omesh.mesh2d(nodes, cells)
    points = np.array([[0.3, -0.2], [2.5, -1.7], [5.9, -3.9], [3.0, -2.0]])
    ids, w = omesh.locate2d(m, points)
    assert np.allclose(np.einsum("pk,pkd->pd", w, nodes[cells[ids]]), points)
    with pytest.raises(ValueError):
        omesh.locate2d(m, np.array([[50.0, 50.0]]))


def test_unit_tetrahedron_and_box_faces():
    m = omesh.mesh3d(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]]), [[0, 1, 2, 3]])
    assert math.isclose(m["volumes"][0], 1 / 6)
    nodes, tets = adt.cases.box_mesh()
    box = omesh.mesh3d(nodes, tets)
    assert box["boundary_faces"].shape[0] == 2 * 2 * (6 * 6 + 6 * 4 + 6 * 4)
    assert math.isclose(box["volumes"].sum(), 12 * 12 * 8)
    assert math.isclose(box["boundary_face_areas"][box["surface_face_mask"]].sum(), 144.0)


# ---------------------------------------------------------------- element templates


def _duffy_triangle(order=8):
    """Exact-for-polynomials rule on the reference triangle via a collapsed tensor Gauss rule."""

    x, w = np.polynomial.legendre.leggauss(order)
    x, w = 0.5 * (x + 1), 0.5 * w
    u, v = np.meshgrid(x, x, indexing="ij")
    wu, wv = np.meshgrid(w, w, indexing="ij")
    pts = np.column_stack((u.ravel(), (v * (1 - u)).ravel()))
    return pts, (wu * wv * (1 - u)).ravel()


def test_p1_triangle_closed_form():
    nodes, cells = adt.cases.grid_triangles(3, 2)
    K, M = fem.p1_triangle(nodes.astype(float), cells)
    for c in range(cells.shape[0]):
        p = nodes[cells[c]]
        e1, e2 = p[1] - p[0], p[2] - p[0]
        area = 0.5 * abs(e1[0] * e2[1] - e1[1] * e2[0])
        b = np.array([p[1, 1] - p[2, 1], p[2, 1] - p[0, 1], p[0, 1] - p[1, 1]])
        cc = np.array([p[2, 0] - p[1, 0], p[0, 0] - p[2, 0], p[1, 0] - p[0, 0]])
        assert np.allclose(K[c], (np.outer(b, b) + np.outer(cc, cc)) / (4 * area), atol=1e-14)
        assert np.allclose(M[c], area / 12 * (np.ones((3, 3)) + np.eye(3)), atol=1e-14)


def test_p2_triangle_templates_match_an_independent_rule():
    nodes = np.array([[0.0, 0.0], [2.0, 0.3], [0.4, 1.5]])
    K, M = fem.p2_triangle(nodes, np.array([[0, 1, 2]]))
    pts, wts = _duffy_triangle()
    jac = np.column_stack((nodes[1] - nodes[0], nodes[2] - nodes[0]))
    det = abs(np.linalg.det(jac))
    shape = fem.p2_triangle_shape(pts)
    h = 1e-6
    grad_ref = np.stack([(fem.p2_triangle_shape(pts + [h, 0]) - fem.p2_triangle_shape(pts - [h, 0])) / (2 * h),
                         (fem.p2_triangle_shape(pts + [0, h]) - fem.p2_triangle_shape(pts - [0, h])) / (2 * h)], -1)
    grad = np.einsum("qia,ab->qib", grad_ref, np.linalg.inv(jac))
    assert np.allclose(M[0], det * np.einsum("q,qi,qj->ij", wts, shape, shape), atol=1e-12)
    assert np.allclose(K[0], det * np.einsum("q,qid,qjd->ij", wts, grad, grad), atol=1e-7)


def test_q1_rectangle_closed_form():
    a, b = 2.0, 0.5
    nodes = np.array([[0, 0], [a, 0], [a, b], [0, b]], dtype=float)
    K, M = fem.q1_quad(nodes, np.array([[0, 1, 2, 3]]))
    kx = np.array([[2, -2, -1, 1], [-2, 2, 1, -1], [-1, 1, 2, -2], [1, -1, -2, 2]]) / 6.0
    ky = np.array([[2, 1, -1, -2], [1, 2, -2, -1], [-1, -2, 2, 1], [-2, -1, 1, 2]]) / 6.0
    assert np.allclose(K[0], b / a * kx + a / b * ky, atol=1e-14)
    assert np.allclose(M[0], a * b / 36 * np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]]), atol=1e-14)


def test_tet_templates_invariants_and_p2_against_an_independent_rule():
    nodes = np.array([[0.0, 0.0, 0.0], [1.2, 0.1, 0.0], [0.2, 0.9, 0.1], [0.1, 0.3, 1.1]])
    cells = np.array([[0, 1, 2, 3]])
    _, vol, K1 = fem.tet_p1(nodes, cells)
    assert np.allclose(K1[0].sum(axis=1), 0.0, atol=1e-14) and np.allclose(K1[0], K1[0].T)
    K2 = fem.tet_p2_stiffness(nodes, cells)
    # independent: collapsed tensor Gauss rule on the reference tet, FD gradients of the shape functions
    x, w = np.polynomial.legendre.leggauss(6)
    x, w = 0.5 * (x + 1), 0.5 * w
    U, V, W_ = np.meshgrid(x, x, x, indexing="ij")
    WU, WV, WW = np.meshgrid(w, w, w, indexing="ij")
    xi = U.ravel()
    eta = (V * (1 - U)).ravel()
    zeta = (W_ * (1 - U) * (1 - V)).ravel()
    wt = (WU * WV * WW * (1 - U) ** 2 * (1 - V)).ravel()
    ref = np.column_stack((xi, eta, zeta))

    def shape(p):
        return fem.tet_p2_shape(np.column_stack((1 - p.sum(axis=1), p)))

    h = 1e-6
    g = np.stack([(shape(ref + h * e) - shape(ref - h * e)) / (2 * h) for e in np.eye(3)], -1)
    jac = np.stack([nodes[i] - nodes[0] for i in (1, 2, 3)], axis=-1)
    phys = np.einsum("qia,ab->qib", g, np.linalg.inv(jac))
    expected = abs(np.linalg.det(jac)) * np.einsum("q,qid,qjd->ij", wt, phys, phys)
    assert np.allclose(K2[0], expected, atol=1e-6)
    assert np.allclose(K2[0].sum(axis=1), 0.0, atol=1e-12)


def test_robin_2p5d_coefficient_is_the_log_derivative_of_the_image_field():
    nodes, cells = adt.cases.grid_triangles(12, 6)
    m = omesh.mesh2d(nodes, cells)
    src = np.array([5.0, 0.0])
    k = 0.3
    alpha = fem.robin_2p5d(m, src, k)
    far = ~m["surface_edge_mask"]
    assert np.all(alpha[m["surface_edge_mask"]] == 0.0) and np.all(alpha[far] > 0)
    image = np.array([src[0], 2 * m["surface_level"] - src[1]])

    def u(x):
        return K0(k * np.linalg.norm(x - src, axis=-1)) + K0(k * np.linalg.norm(x - image, axis=-1))

    c, n = m["boundary_edge_centers"][far], m["boundary_edge_normals"][far]
    h = 1e-6
    dudn = (u(c + h * n) - u(c - h * n)) / (2 * h)
    np.testing.assert_allclose(alpha[far], -dudn / u(c), rtol=1e-6)


def test_robin_3d_coefficient_is_the_log_derivative_of_the_point_source_field():
    """``alpha = max(r . n / r^2, 0)`` is what makes ``u = 1/|x - c|`` satisfy
    ``du/dn + alpha u = 0`` on the far faces: it equals ``-d log u / dn``, checked by central
    differences of the point-source field along each face normal (a sign or radius mistake
    fails).  Faces whose outward normal points back towards the centre are clipped to 0."""

    nodes, tets = adt.cases.box_mesh(relief=1.0)
    m = omesh.mesh3d(nodes, tets)
    center = np.array([5.5, 6.0, m["boundary_face_centers"][m["surface_face_mask"], 2].max()])
    far = ~m["surface_face_mask"]
    c, n = m["boundary_face_centers"][far], m["boundary_face_normals"][far]
    alpha = fem.robin_coefficient_3d(c, n, center)

    def u(x):
        return 1.0 / np.linalg.norm(x - center, axis=-1)

    h = 1e-6
    dudn = (u(c + h * n) - u(c - h * n)) / (2 * h)
    expected = np.maximum(-dudn / u(c), 0.0)
    np.testing.assert_allclose(alpha, expected, rtol=1e-6, atol=1e-12)
    assert np.all(alpha >= 0) and np.count_nonzero(alpha) > 0.9 * alpha.size
    # a face centre inward of the centre (normal towards it) gets no Robin term
    assert fem.robin_coefficient_3d(np.array([[0.0, 0.0, -1.0]]), np.array([[0.0, 0.0, 1.0]]), np.zeros(3))[0] == 0.0
    assert fem.robin_coefficient_3d(np.zeros((1, 3)), np.array([[0.0, 0.0, 1.0]]), np.zeros(3))[0] == 0.0


def test_p2_face_mass_matches_an_independent_rule():
    pts, wts = _duffy_triangle()
    bary = np.column_stack((1.0 - pts.sum(axis=1), pts))
    edges = np.stack([4.0 * bary[:, i] * bary[:, j] for i, j in ((0, 1), (1, 2), (2, 0))], -1)
    shape = np.concatenate((bary * (2.0 * bary - 1.0), edges), axis=1)
    area = 0.37
    expected = 2.0 * area * np.einsum("q,qi,qj->ij", wts, shape, shape)  # reference triangle has area 1/2
    assert relmax(fem.p2_face_mass(area), expected) < 1e-13
    p1 = 2.0 * area * np.einsum("q,qi,qj->ij", wts, bary, bary)
    assert relmax(area * fem.P1_FACE_MASS, p1) < 1e-14


@pytest.mark.parametrize("order", [1, 2])
def test_robin_3d_faces_enter_the_templates_as_a_surface_integral(order):
    """Every cell's Robin term equals ``sum_f alpha_f int_f phi_i phi_j dS`` over its far faces,
    with the tetrahedron's own shape functions evaluated at points of the face (barycentric
    coordinates solved from the vertex coordinates, a collapsed Gauss rule on the face): this
    pins the face mass, its area scaling and the face-to-cell local DOF mapping together."""

    nodes, tets = adt.cases.box_mesh(relief=1.0)
    quads = osurvey.wenner_alpha(5)
    line = np.flatnonzero(np.isclose(nodes[:, 1], 6.0))
    top = [line[np.isclose(nodes[line, 0], x)][np.argmax(nodes[line[np.isclose(nodes[line, 0], x)], 2])]
           for x in np.unique(nodes[line, 0])[1:6]]  # the surface node of five columns
    f = ert3d.Forward3D(nodes, tets, nodes[top], quads, element_order=order)
    m = f.mesh
    far = np.flatnonzero(~m["surface_face_mask"])
    alpha = fem.robin_coefficient_3d(m["boundary_face_centers"][far], m["boundary_face_normals"][far],
                                     f.positions.mean(axis=0))
    pts, wts = _duffy_triangle(6)
    tri = np.column_stack((1.0 - pts.sum(axis=1), pts))
    expected = np.zeros_like(f.cell_t)
    for face, a in zip(far, alpha):
        cell = int(m["boundary_face_cells"][face])
        corners = m["nodes"][m["boundary_faces"][face]]
        x = tri @ corners  # points on the face
        verts = m["nodes"][m["cells"][cell]]
        jac = np.column_stack([verts[i] - verts[0] for i in (1, 2, 3)])
        lam = np.linalg.solve(jac, (x - verts[0]).T).T
        bary = np.column_stack((1.0 - lam.sum(axis=1), lam))
        shape = fem.tet_p2_shape(bary) if order == 2 else bary
        area = 0.5 * np.linalg.norm(np.cross(corners[1] - corners[0], corners[2] - corners[0]))
        expected[cell] += a * 2.0 * area * np.einsum("q,qi,qj->ij", wts, shape, shape)
    robin = f.cell_t - f.volume_t
    assert relmax(robin, expected) < 1e-12
    assert np.count_nonzero(np.abs(robin).sum(axis=(1, 2))) == np.unique(m["boundary_face_cells"][far[alpha > 0]]).size


# ---------------------------------------------------------------- operators


@pytest.fixture(scope="module")
def quad_oracle():
    c = adt.cases.quad_strip_case()
    return c, ert2p5d.Forward2p5D(c.nodes, c.cells, c.electrodes, c.measurements, c.surface_node_ids)


def test_assembly_equals_a_dense_brute_force_sum(quad_oracle):
    c, f = quad_oracle
    w = 3
    dense = np.zeros((f.N, f.N))
    cells = f.mesh["cells"]
    for e in range(f.C):
        dense[np.ix_(cells[e], cells[e])] += c.conductivity[e] * (f.stiffness[e] + f.k[w] ** 2 * f.mass[e])
    for b in range(f.bdofs.shape[0]):
        dense[np.ix_(f.bdofs[b], f.bdofs[b])] += c.conductivity[f.bcells[b]] * f.alpha[w, b] * f.bmass[b]
    assert relmax(f.matrix(c.conductivity, w).toarray(), dense) < 1e-14


def test_2p5d_homogeneous_scaling_reciprocity_and_polarity(quad_oracle):
    c, f = quad_oracle
    assert relmax(f.apparent_resistivity(np.full(f.C, 0.01)), 100.0) < 2e-3
    base = f.apparent_resistivity(c.conductivity)
    assert relmax(f.apparent_resistivity(c.conductivity / 3), 3 * base) < 1e-9
    swapped = ert2p5d.Forward2p5D(c.nodes, c.cells, c.electrodes, c.measurements[:, [2, 3, 0, 1]], c.surface_node_ids)
    assert relmax(swapped.resistance(c.conductivity), f.resistance(c.conductivity)) < 1e-8
    flipped = c.measurements.copy()
    flipped[::2, 2], flipped[::2, 3] = c.measurements[::2, 3], c.measurements[::2, 2]
    pol = ert2p5d.Forward2p5D(c.nodes, c.cells, c.electrodes, flipped, c.surface_node_ids)
    assert relmax(pol.apparent_resistivity(c.conductivity), base) < 1e-10
    assert relmax(pol.resistance(c.conductivity)[::2], -f.resistance(c.conductivity)[::2]) < 1e-10


def test_2p5d_derivatives_dot_tests_fd_and_identities(quad_oracle):
    c, f = quad_oracle
    s = c.conductivity
    rng = np.random.default_rng(1)
    d, g = rng.standard_normal(f.C) * 1e-3, rng.standard_normal(f.D)
    assert (f.jvp(s, d) * g).sum() == pytest.approx((d * f.vjp(s, g)).sum(), rel=1e-9)
    assert (f.normal_jvp(s, d) * g).sum() == pytest.approx((d * f.normal_vjp(s, g)).sum(), rel=1e-9)
    direction = rng.standard_normal(f.C) * s
    h = 1e-5
    numeric = (f.resistance(s + h * direction) - f.resistance(s - h * direction)) / (2 * h)
    assert relmax(f.jvp(s, direction), numeric) < 1e-5
    J = f.jacobian(s)
    assert relmax(J.T @ g, f.normal_vjp(s, g)) < 1e-10 and relmax(J @ d, f.normal_jvp(s, d)) < 1e-10
    exact = f.exact_jacobian(s)
    assert relmax(exact @ d, f.jvp(s, d)) < 1e-9
    log_sums = -(exact * s[None]).sum(axis=1) / f.resistance(s)
    assert relmax(log_sums, np.ones(f.D)) < 1e-6


def test_sensitivity_equals_a_looped_brute_force():
    nodes, cells = adt.cases.grid_triangles(8, 3)
    electrodes = np.column_stack((np.arange(1.0, 8.0), np.zeros(7)))
    quads = adt.cases.dipole_dipole(7, spacings=(1,), limit=6)
    f = ert2p5d.Forward2p5D(nodes, cells, electrodes, quads)
    s = np.random.default_rng(2).uniform(0.005, 0.05, f.C)
    J = f.jacobian(s)
    phi = np.concatenate((f.fields(s)[0], np.zeros((f.W, 1, f.N))), axis=1)
    loop = np.zeros_like(J)
    for d, (a, b, m, n) in enumerate(f.slots):
        for c in range(f.C):
            dofs = cells[c]
            for w in range(f.W):
                T = f.q[w] * (f.stiffness[c] + f.k[w] ** 2 * f.mass[c])
                loop[d, c] -= (phi[w, m, dofs] - phi[w, n, dofs]) @ T @ (phi[w, a, dofs] - phi[w, b, dofs])
    assert relmax(J, loop) < 1e-13


@pytest.fixture(scope="module")
def box():
    return adt.cases.box_case()


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("boundary", ["mixed", "dirichlet"])
def test_3d_homogeneous_half_space_is_exact(box, order, boundary):
    f = ert3d.Forward3D(box.nodes, box.cells, box.electrodes, box.measurements, boundary_mode=boundary,
                        element_order=order)
    assert relmax(f.apparent_resistivity(np.full(f.C, 0.01)), 100.0) < 1e-12


def test_3d_jacobian_fd_dot_tests_poles_and_orders(box):
    f = ert3d.Forward3D(box.nodes, box.cells, box.electrodes, box.measurements)
    s = box.conductivity
    rng = np.random.default_rng(0)
    direction = rng.standard_normal(f.C) * 1e-3 * s
    h = 1e-3
    numeric = (f.resistance(s + h * direction) - f.resistance(s - h * direction)) / (2 * h)
    J = f.jacobian(s, robin=True)
    assert np.abs(J @ direction - numeric).max() / np.abs(numeric).max() < 1e-5
    d, g = rng.standard_normal(f.C) * 1e-3, rng.standard_normal(f.D)
    assert (f.jvp(s, d) * g).sum() == pytest.approx((d * f.vjp(s, g)).sum(), rel=1e-9)
    assert (f.normal_jvp(s, d) * g).sum() == pytest.approx((d * f.normal_vjp(s, g)).sum(), rel=1e-9)
    assert relmax(J.T @ g, f.vjp(s, g)) < 1e-10
    a, b, m, n = box.measurements.T
    poles = np.vstack((np.column_stack((a, np.full_like(a, R), m, n)), np.column_stack((b, np.full_like(b, R), m, n))))
    pole = ert3d.Forward3D(box.nodes, box.cells, box.electrodes, poles)
    rp = pole.resistance(s)
    D = len(a)
    assert relmax(f.resistance(s), rp[:D] - rp[D:]) < 1e-10
    jp = pole.jacobian(s)
    assert relmax(f.jacobian(s), jp[:D] - jp[D:]) < 1e-10
    p2 = ert3d.Forward3D(box.nodes, box.cells, box.electrodes, box.measurements, element_order=2)
    r1, r2 = f.apparent_resistivity(s), p2.apparent_resistivity(s)
    assert np.allclose(r2, r1, rtol=0.15) and not np.allclose(r2, r1, rtol=1e-9)


def test_damping_network_oracle_backward_matches_central_differences(monkeypatch):
    from oracle import gnn
    from tests.test_gpu_kernels import random_spd_batch

    rng = np.random.default_rng(0)
    indptr, indices, values, _ = random_spd_batch(rng, side=6, S=1)
    flags = (rng.integers(0, 2, indices.size) << 1).astype(np.int64)
    flat = rng.standard_normal(gnn.PARAM_COUNT) * 0.2
    monkeypatch.setattr(gnn, "bf16_round", lambda x: np.asarray(x, dtype=np.float64))
    omega, cache = gnn.forward(flat, indptr, indices, values[0], flags, 0.6)
    domega = rng.standard_normal(omega.size)
    grad = gnn.backward(cache, domega)
    for i in rng.choice(gnn.PARAM_COUNT, 60, replace=False):
        e = np.zeros_like(flat)
        e[i] = 1e-6
        fp = gnn.forward(flat + e, indptr, indices, values[0], flags, 0.6)[0] @ domega
        fm = gnn.forward(flat - e, indptr, indices, values[0], flags, 0.6)[0] @ domega
        assert abs((fp - fm) / 2e-6 - grad[i]) < 1e-7 + 1e-6 * abs(grad[i])


def test_damping_network_oracle_rounds_like_torch_bfloat16_and_is_constant_without_output_layer():
    import torch

    from oracle import gnn

    x = np.random.default_rng(1).standard_normal(10000) * 10.0 ** np.random.default_rng(2).uniform(-8, 8, 10000)
    assert np.array_equal(gnn.bf16_round(x), torch.as_tensor(x, dtype=torch.float32).to(torch.bfloat16).double().numpy())
    from tests.test_gpu_kernels import random_spd_batch

    indptr, indices, values, _ = random_spd_batch(np.random.default_rng(3), side=5, S=1)
    p = gnn.unpack(np.random.default_rng(4).standard_normal(gnn.PARAM_COUNT))
    p["v"][:] = 0.0
    p["d"][:] = 0.0
    omega, _ = gnn.forward(gnn.pack(p), indptr, indices, values[0], np.zeros(indices.size, np.int64), 0.7)
    assert np.all(omega == 0.7)


def test_cpu_jacobi_pcg_matches_the_direct_solve():
    """The CPU baseline's block Jacobi-PCG reaches the tolerance on every column and agrees with
    the sparse LU solution (the 3D oracle's matrix on the test box)."""

    import scipy.sparse.linalg as spla

    case = adt.cases.box_case()
    f = ert3d.Forward3D(case.nodes, case.cells, case.electrodes, case.measurements)
    a = f._assemble_matrix(case.conductivity).tocsr()
    b = np.ascontiguousarray(f.source_rhs[: f.curr_ids.size].T)
    b[:, 1] = 0.0  # a zero right-hand side stays zero
    x, iterations = cpu_pcg.jacobi_pcg(a, b, tol=1e-12, threads=2)
    assert iterations[1] == 0 and np.all(x[:, 1] == 0.0)
    ok = np.linalg.norm(b, axis=0) > 0
    res = np.linalg.norm(b - a @ x, axis=0)[ok] / np.linalg.norm(b, axis=0)[ok]
    assert res.max() <= 1e-12 * 1.01
    direct = spla.splu(a.tocsc()).solve(b)
    assert relmax(x, direct) < 1e-9


def test_cgls_oracle_reaches_least_squares_and_damped_normal_equations():
    """At a tiny tolerance CGLS is the least-squares solution (numpy lstsq) of a full-rank
    overdetermined system; with sqrt(lambda) I rows stacked it is the closed form
    (A^T A + lambda I)^-1 A^T b; b = 0 gives x = 0 after no iteration."""

    import scipy.sparse as sp

    from oracle import cgls as ocgls

    rng = np.random.default_rng(8)
    a = sp.random(60, 25, density=0.3, random_state=8, format="csr") + sp.eye(60, 25, format="csr")
    b = rng.standard_normal(60)
    x, iters = ocgls.cgls(a, b, 500, 1e-28)
    assert relmax(x, np.linalg.lstsq(a.toarray(), b, rcond=None)[0]) < 1e-10 and 0 < iters <= 500
    lam = 0.3
    stacked = sp.vstack((a, np.sqrt(lam) * sp.eye(25, format="csr")), format="csr")
    x2, _ = ocgls.cgls(stacked, np.concatenate((b, np.zeros(25))), 500, 1e-28)
    dense = a.toarray()
    assert relmax(x2, np.linalg.solve(dense.T @ dense + lam * np.eye(25), dense.T @ b)) < 1e-10
    x0, it0 = ocgls.cgls(a, np.zeros(60), 10, 1e-8)
    assert it0 == 0 and np.all(x0 == 0)
    # a loose tolerance stops early, with ||A^T r||^2 below tol ||A^T b||^2
    xl, il = ocgls.cgls(a, b, 500, 1e-2)
    s0, s = a.T @ b, a.T @ (b - a @ xl)
    assert il < iters and float(s @ s) < 1e-2 * float(s0 @ s0)
