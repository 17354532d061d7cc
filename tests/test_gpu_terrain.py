"""2.5D terrain path on the GPU against the FP64 oracle (``forward/ert2p5d.py:170-190, 349-436,
662-671``): the reference's sloped quad-strip case (``tests/conftest.py:27-61``) with the numerical
primary field from serendipity potentials on the H2-refined strips and numerical geometric factors,
a triangle terrain (P1 field, P2 potentials), the reference's homogeneous-terrain and adjoint dot
tests (``tests/test_forward_gpu.py:55-63, 95-115``), and the unit-potential disk cache.  The
oracle receives the same auxiliary discretizations (host-side builders, pinned in
``tests/test_terrain.py``) and computes the unit potentials, fields, factors and derivatives
itself by sparse LU."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ert2p5d as oracle2p5d
from tests.conftest import adt, relmax

pytestmark = pytest.mark.gpu


def arrays(d):
    if d is None:
        return None
    return dict(cell_dofs=d.cell_dofs.numpy(), dof_nodes=np.asarray(d.dof_nodes, float), parents=d.parent_cell_ids.numpy(),
                boundary_dofs=d.boundary_dofs.numpy(), stiffness=d.stiffness.numpy(), mass=d.mass.numpy(),
                boundary_mass=d.boundary_mass.numpy(), alpha=d.boundary_geometries.numpy(),
                electrodes=d.electrode_matrix.numpy(), boundary_cells=d.boundary_cells.numpy())


def build(nodes, cells, electrodes, measurements, surface=None, **options):
    mesh = adt.Mesh.from_arrays(nodes, cells, surface_node_ids=surface)
    survey = adt.Survey.from_arrays(electrodes, measurements)
    op = adt.ERTForward2p5D.from_mesh_survey(mesh, survey, **options)
    ref = oracle2p5d.Forward2p5D(nodes, cells, electrodes, measurements, surface_node_ids=surface,
                                 wavenumbers=op.wavenumbers.double().numpy(), weights=op.weights.double().numpy(),
                                 terrain=dict(field=arrays(op.discretization),
                                              potential=arrays(op.primary_potential_discretization),
                                              geometric=arrays(op.geometric_discretization)))
    return mesh, op, ref


@pytest.fixture(scope="module")
def strip():
    case = adt.cases.quad_strip_case(slope=0.03)
    mesh, op, ref = build(case.nodes, case.cells, case.electrodes, case.measurements, case.surface_node_ids,
                          topographic_geometric_factor_mode="numerical")
    yield case, mesh, op, ref
    op.close()


def test_terrain_strip_matches_the_oracle(strip):
    case, mesh, op, ref = strip
    assert not mesh.is_flat_surface and op.use_numerical_primary and op.use_numerical_geometric_factors
    assert op.discretization.cell_dofs.shape == (4 * mesh.cell_count, 4)
    assert op.primary_potential_discretization.cell_dofs.shape[1] == 8
    primary = op._unit_primary()
    err_primary = relmax(primary, ref.primary)
    err_gf = relmax(op._geometric_factors().numpy(), ref.geometric_factors())
    sigma = case.conductivity
    rho = op.apparent_resistivity_values(torch.as_tensor(sigma)).numpy()
    err = relmax(rho, ref.apparent_resistivity(sigma))
    print(f"[terrain strip] primary {err_primary:.2e}, geometric factors {err_gf:.2e}, rhoa {err:.2e}")
    assert err_primary < 1e-8 and err_gf < 1e-8 and err < 1e-6


def test_homogeneous_terrain_with_numerical_geometric_factors(strip):
    _, mesh, op, _ = strip
    rho = op.apparent_resistivity_values(torch.full((mesh.cell_count,), 0.01, dtype=torch.float64)).numpy()
    assert relmax(rho, np.full_like(rho, 100.0)) < 2e-2  # the reference's bound (test_forward_gpu.py:55-63)


def test_terrain_derivatives_match_the_oracle_and_pass_dot_tests(strip):
    case, mesh, op, ref = strip
    sigma = case.conductivity
    s = torch.as_tensor(sigma)
    jac = op.jacobian(s).numpy()
    err_j = relmax(jac, ref.jacobian(sigma))
    rng = np.random.default_rng(4)
    g, d = rng.standard_normal(ref.D), rng.standard_normal(ref.C) * 1e-3
    vjp, jvp = op.vjp(s, torch.as_tensor(g)).numpy(), op.jvp(s, torch.as_tensor(d)).numpy()
    err_v, err_t = relmax(vjp, ref.vjp(sigma, g)), relmax(jvp, ref.jvp(sigma, d))
    nv, nj = op.normal_vjp(s, torch.as_tensor(g)).numpy(), op.normal_jvp(s, torch.as_tensor(d)).numpy()
    print(f"[terrain strip] J {err_j:.2e}, vjp {err_v:.2e}, jvp {err_t:.2e}, "
          f"normal vjp {relmax(nv, ref.normal_vjp(sigma, g)):.2e}, normal jvp {relmax(nj, ref.normal_jvp(sigma, d)):.2e}")
    assert err_j < 1e-5 and err_v < 1e-5 and err_t < 1e-5
    assert relmax(nv, ref.normal_vjp(sigma, g)) < 1e-5 and relmax(nj, ref.normal_jvp(sigma, d)) < 1e-5
    assert float(jvp @ g) == pytest.approx(float(d @ vjp), rel=1e-6)  # test_forward_gpu.py:95-115
    assert float(nj @ g) == pytest.approx(float(d @ nv), rel=1e-6)
    assert relmax(jac @ d, nj) < 1e-8  # J d = normal_jvp


def test_triangle_terrain_uses_p1_fields_and_p2_potentials():
    nodes, cells = adt.cases.grid_triangles(nx=8, ny=4, slope=0.1)
    top = np.flatnonzero(np.isclose(nodes[:, 1], 0.1 * nodes[:, 0]))
    electrodes = nodes[top[np.argsort(nodes[top, 0])]]
    measurements = adt.cases.wenner_alpha(electrodes.shape[0])
    mesh, op, ref = build(nodes, cells, electrodes, measurements, topographic_geometric_factor_mode="numerical")
    assert op.discretization.cell_dofs.shape == (4 * mesh.cell_count, 3)
    assert op.primary_potential_discretization.cell_dofs.shape[1] == 6
    sigma = np.exp(np.random.default_rng(5).uniform(-1.0, 1.0, mesh.cell_count)) * 0.01
    rho = op.apparent_resistivity_values(torch.as_tensor(sigma)).numpy()
    err = relmax(rho, ref.apparent_resistivity(sigma))
    print(f"[triangle terrain] rhoa {err:.2e}")
    assert err < 1e-6
    op.close()


def test_terrain_unit_potentials_come_back_from_the_disk_cache(tmp_path):
    case = adt.cases.quad_strip_case(nx=14, nz=5, n_electrodes=8, slope=0.05)
    options = dict(topographic_geometric_factor_mode="numerical", terrain_cache_dir=tmp_path)
    mesh, first, _ = build(case.nodes, case.cells, case.electrodes, case.measurements, case.surface_node_ids, **options)
    rho = first.apparent_resistivity_values(torch.as_tensor(case.conductivity)).numpy()
    files = sorted(p.name for p in tmp_path.iterdir())
    assert len(files) >= 3  # field primary, potential and geometric unit potentials
    second = adt.ERTForward2p5D.from_mesh_survey(mesh, adt.Survey.from_arrays(case.electrodes, case.measurements),
                                                 **options)
    again = second.apparent_resistivity_values(torch.as_tensor(case.conductivity)).numpy()
    assert np.array_equal(again, rho) and sorted(p.name for p in tmp_path.iterdir()) == files
    first.close()
    second.close()


R = adt.REMOTE_ELECTRODE


def with_poles(quads, count):
    """Dipole-dipole rows plus pole-dipole, dipole-pole and pole-pole rows (``test_remote_electrodes_gpu.py:20-31``)."""

    poles = [(a, R, a + 1 + s, a + 2 + s) for a in range(count - 3) for s in (0, 1) if a + 2 + s < count]
    poles += [(a, a + 1, a + 2 + s, R) for a in range(count - 2) for s in (0, 1) if a + 2 + s < count]
    poles += [(a, R, a + 1 + s, R) for a in range(count - 1) for s in (0, 1) if a + 1 + s < count]
    return np.vstack((np.asarray(quads), np.asarray(poles)))


def test_terrain_spaces_have_robin_far_field_boundaries_and_sane_pole_pole_readings():
    """``test_remote_electrodes_gpu.py:140-175``: the sides and bottom of both auxiliary spaces
    carry the 2.5D mixed condition and the ground surface does not; homogeneous pole-pole
    resistances on the sloped strip stay within half of the flat strip's (natural outer
    boundaries would make them orders of magnitude larger).  The terrain resistances also match
    the oracle, remote rows included."""

    def pole_pole(slope):
        case = adt.cases.quad_strip_case(anomaly=False, slope=slope)
        quads = with_poles(case.measurements, case.electrodes.shape[0])
        mesh = adt.Mesh.from_arrays(case.nodes, case.cells, surface_node_ids=case.surface_node_ids)
        op = adt.ERTForward2p5D.from_mesh_survey(mesh, adt.Survey.from_arrays(case.electrodes, quads))
        sigma = np.full(mesh.cell_count, 0.01)
        resistance = op.resistance(torch.as_tensor(sigma)).numpy()
        return case, quads, op, sigma, resistance

    _, quads, flat_op, _, flat = pole_pole(0.0)
    case, _, op, sigma, terrain = pole_pole(0.03)
    for d in (op.discretization, op.primary_potential_discretization):
        surface = np.asarray(d.geometry.surface_edge_mask, dtype=bool)
        robin = np.asarray(d.boundary_geometries)
        assert surface.any() and (~surface).any()
        assert np.all(robin[:, surface] == 0) and np.all(np.abs(robin[:, ~surface]).max(axis=0) > 0)
    rows = (quads[:, 1] == R) & (quads[:, 3] == R)
    assert rows.sum() > 0 and np.all(np.abs(terrain[rows] - flat[rows]) < 0.5 * flat[rows])
    ref = oracle2p5d.Forward2p5D(case.nodes, case.cells, case.electrodes, quads, surface_node_ids=case.surface_node_ids,
                                 wavenumbers=op.wavenumbers.double().numpy(), weights=op.weights.double().numpy(),
                                 terrain=dict(field=arrays(op.discretization),
                                              potential=arrays(op.primary_potential_discretization),
                                              geometric=arrays(op.geometric_discretization)))
    err = relmax(terrain, ref.resistance(sigma))
    print(f"[terrain poles] resistance vs oracle {err:.2e}")
    assert err < 1e-6
    flat_op.close()
    op.close()


def test_swapped_receivers_on_terrain_give_the_same_data_and_derivatives(strip):
    """``test_polarity_gpu.py:53-90`` (terrain case): reversed receivers flip the sign of the
    numerical geometric factor and of the resistance, and leave the apparent resistivities and
    the gradient of their logarithm unchanged."""

    case, mesh, op, _ = strip
    quads = case.measurements.copy()
    rows = np.arange(0, len(quads), 2)
    quads[rows, 2], quads[rows, 3] = case.measurements[rows, 3], case.measurements[rows, 2]
    flipped = adt.ERTForward2p5D.from_mesh_survey(mesh, adt.Survey.from_arrays(case.electrodes, quads),
                                                  topographic_geometric_factor_mode="numerical")
    sigma = torch.as_tensor(case.conductivity)
    assert bool(torch.all(flipped._geometric_factors()[rows] < 0))
    rho = flipped.apparent_resistivity_values(sigma)
    assert bool(torch.all(rho > 0)) and relmax(rho, op.apparent_resistivity_values(sigma)) < 1e-10
    assert relmax(flipped.resistance(sigma)[rows], -op.resistance(sigma)[rows]) < 1e-10
    grads = []
    for forward in (op, flipped):
        model = sigma.clone().requires_grad_()
        torch.log(adt.apparent_resistivity_autograd(model, forward)).sum().backward()
        grads.append(model.grad)
    assert relmax(grads[1], grads[0]) < 1e-8
    flipped.close()
