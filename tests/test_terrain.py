"""Host-side builders of the 2.5D terrain path (``forward/discretization.py:159-592``), each pinned
against what the mathematics fixes: shape functions by the Kronecker property and finite
differences, element templates by exact integrals of quadratic functions (and the P2 triangle
against the oracle's independently pinned template), refinement by areas and parents, the
quadratic topologies by midpoint coordinates, the electrode interpolation by reproducing linear
and quadratic functions along the surface, and the columnar expansion by recovering the strip
quads it was cut from."""

from __future__ import annotations

import importlib

import numpy as np
import pytest

from oracle import fem as ofem
from tests.conftest import adt

terrain = importlib.import_module("ad-tlert_amd.terrain")
fem = importlib.import_module("ad-tlert_amd.fem")

CORNERS = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])
SERENDIPITY_NODES = np.vstack((CORNERS, 0.5 * (CORNERS + np.roll(CORNERS, -1, axis=0))))


def strip(slope=0.03, nx=12, nz=5, electrodes=7):
    c = adt.cases.quad_strip_case(nx=nx, nz=nz, n_electrodes=electrodes, slope=slope)
    return c, adt.Mesh.from_arrays(c.nodes, c.cells, surface_node_ids=c.surface_node_ids)


def test_serendipity_shape_functions_are_nodal_and_their_gradients_differentiate_them():
    values, grads = fem.serendipity_shape(SERENDIPITY_NODES)
    assert np.allclose(values, np.eye(8), atol=1e-15)
    rng = np.random.default_rng(0)
    pts = rng.uniform(-0.9, 0.9, (20, 2))
    v, g = fem.serendipity_shape(pts)
    assert np.allclose(v.sum(axis=1), 1.0) and np.allclose(g.sum(axis=1), 0.0)
    h = 1e-6
    for a in range(2):
        e = np.zeros(2)
        e[a] = h
        fd = (fem.serendipity_shape(pts + e)[0] - fem.serendipity_shape(pts - e)[0]) / (2 * h)
        assert np.allclose(fd, g[..., a], atol=1e-8)


def _rectangle(a, b, x0=0.3, y0=-1.0):
    nodes = np.array([[x0, y0], [x0 + a, y0], [x0 + a, y0 + b], [x0, y0 + b]])
    return nodes, np.array([[0, 1, 2, 3]])


def test_serendipity_templates_integrate_quadratics_exactly():
    a, b, x0, y0 = 1.7, 0.6, 0.3, -1.0
    nodes, cells = _rectangle(a, b, x0, y0)
    k, m = fem.serendipity_quad(nodes, cells)
    pts = np.array([x0, y0]) + (SERENDIPITY_NODES + 1.0) / 2.0 * np.array([a, b])  # positions of the 8 DOFs
    x, y = pts[:, 0] - x0, pts[:, 1] - y0
    u = x * x  # in the serendipity space
    assert np.isclose(u @ k[0] @ u, 4 * a ** 3 * b / 3)  # int |grad x^2|^2
    assert np.isclose(x @ m[0] @ x, a ** 3 * b / 3) and np.isclose(m[0].sum(), a * b)
    v = x * y
    assert np.isclose(v @ k[0] @ v, (b ** 3 * a + a ** 3 * b) / 3)  # int y^2 + x^2
    assert np.allclose(k[0] @ np.ones(8), 0.0) and np.allclose(k[0], k[0].T)


def test_p2_triangle_templates_equal_the_oracle_rule():
    nodes = np.array([[0.0, 0.0], [1.3, 0.2], [0.4, 0.9], [1.6, 1.1]])
    cells = np.array([[0, 1, 2], [1, 3, 2]])
    k, m = fem.p2_triangle(nodes, cells)
    ko, mo = ofem.p2_triangle(nodes, cells)
    assert np.allclose(k, ko, rtol=1e-13, atol=1e-14) and np.allclose(m, mo, rtol=1e-13, atol=1e-14)


def test_quadratic_edge_mass_integrates_products_of_quadratics():
    length = 2.5
    mb = fem.quadratic_edge_mass(np.array([length]))[0]
    u = np.array([0.0, (length / 2) ** 2, length ** 2])  # t^2 at (end, midpoint, end)
    assert np.isclose(u @ mb @ u, length ** 5 / 5) and np.isclose(mb.sum(), length)


@pytest.mark.parametrize("quads", [True, False])
def test_refinement_keeps_area_parents_and_the_surface(quads):
    if quads:
        _, mesh = strip()
        fine, parents = terrain.refine_quads(mesh)
    else:
        mesh = adt.Mesh.from_arrays(*adt.cases.grid_triangles(nx=6, ny=4, slope=0.1))
        fine, parents = terrain.refine_triangles(mesh)
    assert fine.cell_count == 4 * mesh.cell_count and np.array_equal(np.bincount(parents), np.full(mesh.cell_count, 4))
    area = np.asarray(mesh.cell_areas, float)
    assert np.allclose(np.bincount(parents, weights=np.asarray(fine.cell_areas, float)), area)
    surface = np.asarray(fine.surface_nodes, float)
    assert surface.shape[0] == 2 * np.asarray(mesh.surface_nodes).shape[0] - 1
    assert np.allclose(surface[1::2], 0.5 * (surface[:-1:2] + surface[2::2]))  # inserted midpoints


@pytest.mark.parametrize("quads", [True, False])
def test_quadratic_topologies_place_midpoints(quads):
    if quads:
        _, mesh = strip()
        nodes, cells, bnd = terrain.serendipity_topology(mesh)
        corners, mids = cells[:, :4], cells[:, 4:]
        pairs = np.stack((corners, np.roll(corners, -1, axis=1)), axis=-1)
    else:
        mesh = adt.Mesh.from_arrays(*adt.cases.grid_triangles(nx=5, ny=3, slope=0.1))
        nodes, cells, bnd = terrain.p2_topology(mesh)
        corners, mids = cells[:, :3], cells[:, 3:]
        pairs = np.stack((corners, np.roll(corners, -1, axis=1)), axis=-1)
        bnd = bnd[:, [0, 2, 1]]  # (a, b, m) -> (a, m, b)
    assert np.allclose(nodes[mids], 0.5 * (nodes[pairs[..., 0]] + nodes[pairs[..., 1]]))
    assert np.allclose(nodes[bnd[:, 1]], 0.5 * (nodes[bnd[:, 0]] + nodes[bnd[:, 2]]))
    edges = np.unique(np.sort(pairs.reshape(-1, 2), axis=1), axis=0)
    assert nodes.shape[0] == mesh.node_count + edges.shape[0]


def test_surface_interpolation_reproduces_linear_and_quadratic_functions():
    c, mesh = strip(electrodes=7)
    fine, _ = terrain.refine_quads(mesh)
    rng = np.random.default_rng(3)
    surface = np.asarray(fine.surface_nodes, float)
    seg = rng.integers(0, surface.shape[0] - 1, 9)
    t = rng.uniform(0.05, 0.95, 9)
    pts = surface[seg] + t[:, None] * (surface[seg + 1] - surface[seg])
    nodes = np.asarray(fine.nodes, float)
    rows = terrain.surface_interpolation(fine, None, nodes.shape[0], pts)
    f = 2.0 * nodes[:, 0] - 3.0 * nodes[:, 1]
    assert np.allclose(rows @ f, 2.0 * pts[:, 0] - 3.0 * pts[:, 1])
    qn, _, qb = terrain.serendipity_topology(fine)
    rows = terrain.surface_interpolation(fine, qb, qn.shape[0], pts)
    # a quadratic in the arc parameter of each straight surface segment
    ids = np.asarray(fine.surface_node_ids)
    lookup = {tuple(sorted((a, b))): m for a, m, b in qb.tolist()}
    for i in range(pts.shape[0]):
        a, b = int(ids[seg[i]]), int(ids[seg[i] + 1])
        g = np.zeros(qn.shape[0])
        g[a], g[lookup[tuple(sorted((a, b)))]], g[b] = 0.0, 0.25, 1.0  # t^2 at (start, middle, end)
        assert np.isclose(rows[i] @ g, t[i] ** 2)
    assert np.allclose(rows.sum(axis=1), 1.0)
    exact = terrain.surface_interpolation(fine, None, nodes.shape[0], surface[:3])
    assert np.array_equal(exact[np.arange(3), ids[:3]], np.ones(3))


def test_columnar_half_cells_expand_back_into_the_strip_quads():
    _, mesh = strip(nx=12, nz=4)
    quads = np.asarray(mesh.cells)
    halves = quads[:, :3]  # one triangle (c0, c1, c2) of every strip quad
    half = adt.Mesh.from_arrays(np.asarray(mesh.nodes), halves, surface_node_ids=np.asarray(mesh.surface_node_ids))
    expanded = terrain.expand_columnar(half)
    assert expanded is not None and expanded[0].cell_count == 2 * mesh.cell_count
    assert np.allclose(np.bincount(expanded[1], weights=np.asarray(expanded[0].cell_areas, float)),
                       np.asarray(mesh.cell_areas, float))
    rebuilt = terrain.structured_quads(half)
    assert rebuilt is not None and rebuilt.cell_count == mesh.cell_count
    # the same corners (nodes unused by the half cells are dropped and re-created: compare positions)
    got = np.sort(np.asarray(rebuilt.nodes, float)[np.asarray(rebuilt.cells)].round(9), axis=1)
    want = np.sort(np.asarray(mesh.nodes, float)[quads].round(9), axis=1)
    assert np.allclose(got, want)
    assert terrain.structured_quads(adt.Mesh.from_arrays(*adt.cases.grid_triangles(6, 4, 0.1)[:2])) is None


def test_build_gives_the_auxiliary_spaces_of_the_reference():
    c, mesh = strip()
    k = np.array([0.05, 0.4])
    field = terrain.build("primary", mesh, c.electrodes, k, refine=True, quadratic=False)
    potential = terrain.build("primary_potential", mesh, c.electrodes, k, refine=True, quadratic=True)
    assert field["cell_dofs"].shape == (4 * mesh.cell_count, 4) and potential["cell_dofs"].shape[1] == 8
    assert np.array_equal(field["parents"], potential["parents"])
    sel = terrain.exact_node_indices(potential["dof_nodes"], field["dof_nodes"])
    assert np.array_equal(sel, np.arange(field["dof_nodes"].shape[0]))  # the field nodes come first
    assert field["alpha"].shape == (2, field["boundary_dofs"].shape[0]) and np.all(field["alpha"] >= 0)
    surface = np.asarray(field["geometry"].surface_edge_mask)
    assert np.all(field["alpha"][:, surface] == 0) and np.all(field["alpha"][:, ~surface] > 0)
    assert np.allclose(field["electrodes"].sum(axis=1), 1.0)
    tri = terrain.build("geometric", adt.Mesh.from_arrays(*adt.cases.grid_triangles(6, 4, 0.1)[:2]),
                        np.array([[1.0, 0.1], [3.0, 0.3]]), k, refine=True, quadratic=True)
    assert tri["cell_dofs"].shape[1] == 6 and tri["boundary_mass"].shape[1:] == (3, 3)
