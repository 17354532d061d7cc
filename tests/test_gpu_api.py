"""The reference operator API on the GPU: autograd bridge, series and field cache, remote
electrodes, polarity, reciprocity, parameter aggregation, numerical geometric factors on 3D
relief, mixed-precision sensitivities, validation and ``close()``.

Mirrors ``tests/test_forward_gpu.py``, ``tests/test_forward3d_gpu.py``,
``tests/test_polarity_gpu.py`` and ``tests/test_remote_electrodes_gpu.py`` of the reference,
with the FP64 oracle as the second opinion where the reference compared two code paths.
"""

import numpy as np
import pytest
import torch

from oracle import ert2p5d as oracle2p5d
from oracle import ert3d as oracle3d
from tests.conftest import adt, relmax

pytestmark = pytest.mark.gpu
R = adt.REMOTE_ELECTRODE


def with_poles(quads, count):
    """Dipole-dipole rows plus pole-dipole, dipole-pole and pole-pole rows (``test_remote_electrodes_gpu.py:20-31``)."""

    poles = [(a, R, a + 1 + s, a + 2 + s) for a in range(count - 3) for s in (0, 1) if a + 2 + s < count]
    poles += [(a, a + 1, a + 2 + s, R) for a in range(count - 2) for s in (0, 1) if a + 2 + s < count]
    poles += [(a, R, a + 1 + s, R) for a in range(count - 1) for s in (0, 1) if a + 1 + s < count]
    return np.vstack((np.asarray(quads), np.asarray(poles)))


def flat_operator(quads=None, **options):
    case = adt.cases.quad_strip_case()
    quads = case.measurements if quads is None else quads
    mesh = adt.Mesh.from_arrays(case.nodes, case.cells, surface_node_ids=case.surface_node_ids)
    survey = adt.Survey.from_arrays(case.electrodes, quads)
    options.setdefault("solver_tolerance", 1e-12)
    return case, adt.ERTForward2p5D.from_mesh_survey(mesh, survey, **options)


def box_operator(quads=None, relief=0.0, **options):
    nodes, tets = adt.cases.box_mesh(relief=relief)
    sigma = np.full(tets.shape[0], 0.01)
    sigma[: tets.shape[0] // 3] = 0.03
    surface = np.isclose(nodes[:, 2], relief * np.sin(nodes[:, 0] / 4.0))
    line = np.flatnonzero(np.isclose(nodes[:, 1], 6.0) & surface)
    line = line[np.argsort(nodes[line, 0])]
    quads = adt.cases.wenner_alpha(len(line)) if quads is None else quads
    mesh = adt.Mesh3D.from_arrays(nodes, tets)
    survey = adt.Survey.from_arrays(nodes[line], quads)
    options.setdefault("solver_tolerance", 1e-12)
    return (nodes, tets, nodes[line], quads, sigma), adt.ERTForward3D.from_mesh_survey(mesh, survey, **options)


def test_autograd_bridge_matches_vjp_and_current_gradient():
    case, forward = flat_operator()
    sigma = torch.as_tensor(case.conductivity).clone().requires_grad_()
    current = torch.tensor(1.5, dtype=torch.float64, requires_grad=True)
    weights = torch.as_tensor(np.random.default_rng(1).standard_normal(forward.survey.measurement_count))
    (adt.apparent_resistivity_autograd(sigma, forward, current) * weights).sum().backward()
    scale = forward._geometric_factors() / 1.5
    assert relmax(sigma.grad, forward.vjp(sigma.detach(), weights * scale)) < 1e-10
    rhoa = adt.apparent_resistivity_autograd(sigma.detach(), forward, current.detach())
    assert float(current.grad) == pytest.approx(-float((weights * rhoa).sum()) / 1.5)
    series = torch.stack([sigma.detach(), sigma.detach() * 1.3]).requires_grad_()
    w2 = torch.stack([weights, 0.7 * weights])
    rho_s = adt.apparent_resistivity_autograd(series, forward)
    (rho_s * w2).sum().backward()
    for t in range(2):
        single = series[t].detach().clone().requires_grad_()
        value = adt.apparent_resistivity_autograd(single, forward)
        (value * w2[t]).sum().backward()
        assert relmax(rho_s[t].detach(), value.detach()) < 1e-10
        assert relmax(series.grad[t], single.grad) < 1e-8


def test_forward_mode_jvp_through_autograd_matches_the_operator():
    case, forward = flat_operator()
    sigma = torch.as_tensor(case.conductivity)
    direction = torch.as_tensor(np.random.default_rng(2).standard_normal(sigma.numel()) * 1e-3)
    import torch.autograd.forward_ad as fwAD

    with fwAD.dual_level():
        dual = fwAD.make_dual(sigma, direction)
        tangent = fwAD.unpack_dual(adt.apparent_resistivity_autograd(dual, forward)).tangent
    expected = forward._geometric_factors() * forward.jvp(sigma, direction)
    assert relmax(tangent, expected) < 1e-10


def test_series_matches_individual_solves_and_uses_the_field_cache():
    case, forward = flat_operator(normal_field_cache_max_entries=8, series_batch_steps=2)
    sigma = torch.as_tensor(case.conductivity)
    series = torch.stack([sigma, sigma * 1.2, sigma * 0.8])
    values = forward.apparent_resistivity_series(series)
    for t in range(3):
        assert relmax(values[t], forward.apparent_resistivity_values(series[t])) < 1e-9
    before = forward.normal_field_cache_info()["hits"]
    forward.normal_vjp_series(series, torch.ones(3, forward.survey.measurement_count, dtype=torch.float64))
    assert forward.normal_field_cache_info()["hits"] >= before + 3
    g = torch.as_tensor(np.random.default_rng(3).standard_normal((3, forward.survey.measurement_count)))
    batched = forward.vjp_series(series, g)
    for t in range(3):
        assert relmax(batched[t], forward.vjp(series[t], g[t])) < 1e-8


@pytest.mark.parametrize("dim", ["2p5d", "3d"])
def test_remote_electrodes_homogeneous_dot_tests_and_jacobian(dim):
    if dim == "2p5d":
        case, _ = flat_operator()
        quads = with_poles(case.measurements, 16)
        case, forward = flat_operator(quads)
        sigma = torch.as_tensor(case.conductivity)
        tol_h = 2e-3
    else:
        (nodes, tets, el, q, s), _ = box_operator()
        quads = with_poles(q, len(el))
        _, forward = box_operator(quads)
        sigma = torch.as_tensor(s)
        tol_h = 1e-6
    homogeneous = forward.apparent_resistivity_values(torch.full((forward.mesh.cell_count,), 0.01))
    assert relmax(homogeneous.numpy(), 100.0) < tol_h
    rng = np.random.default_rng(0)
    direction = torch.as_tensor(rng.standard_normal(sigma.numel()) * 1e-3) * sigma
    cot = torch.as_tensor(rng.standard_normal(forward.survey.measurement_count))
    for jvp, vjp in ((forward.jvp, forward.vjp), (forward.normal_jvp, forward.normal_vjp)):
        assert float((jvp(sigma, direction) * cot).sum()) == pytest.approx(float((direction * vjp(sigma, cot)).sum()),
                                                                            rel=1e-6)
    jac = forward.jacobian(sigma)
    assert relmax(jac.T @ cot, forward.normal_vjp(sigma, cot)) < 1e-8
    assert relmax(jac @ direction, forward.normal_jvp(sigma, direction)) < 1e-8
    h = 1e-5
    d2 = torch.as_tensor(rng.standard_normal(sigma.numel())) * sigma
    numeric = (forward.resistance(sigma + h * d2) - forward.resistance(sigma - h * d2)) / (2 * h)
    assert relmax(forward.jvp(sigma, d2), numeric) < 1e-5


def test_polarity_swap_and_reciprocity():
    case, forward = flat_operator()
    sigma = torch.as_tensor(case.conductivity)
    quads = case.measurements.copy()
    rows = np.arange(0, len(quads), 2)
    quads[rows, 2], quads[rows, 3] = case.measurements[rows, 3], case.measurements[rows, 2]
    _, flipped = flat_operator(quads)
    assert bool(torch.all(flipped.survey.geometric_factors()[rows] < 0))
    assert relmax(flipped.apparent_resistivity_values(sigma), forward.apparent_resistivity_values(sigma)) < 1e-10
    assert relmax(flipped.resistance(sigma)[rows], -forward.resistance(sigma)[rows]) < 1e-10
    grads = []
    for op in (forward, flipped):
        model = sigma.clone().requires_grad_()
        torch.log(adt.apparent_resistivity_autograd(model, op)).sum().backward()
        grads.append(model.grad)
    assert relmax(grads[1], grads[0]) < 1e-8
    _, swapped = flat_operator(case.measurements[:, [2, 3, 0, 1]])
    assert relmax(swapped.resistance(sigma), forward.resistance(sigma)) < 1e-8


@pytest.mark.parametrize("dim", ["2p5d", "3d"])
def test_parameter_aggregation_sums_cells_and_ignores_inactive(dim):
    """Aggregated Jacobian columns (summed inside the Gram kernels) and the aggregated normal vjp
    against the oracle's aggregation of its own per-cell Jacobian (``kernels.py:88-113``):
    inactive cells (-1) are dropped and a column with no cell (``parameter_count`` beyond the
    largest id) is zero."""

    if dim == "2p5d":
        case, forward = flat_operator()
        sigma = case.conductivity
        ref = oracle2p5d.Forward2p5D(case.nodes, case.cells, case.electrodes, case.measurements,
                                     case.surface_node_ids)
    else:
        (nodes, tets, el, quads, sigma), forward = box_operator()
        ref = oracle3d.Forward3D(nodes, tets, el, quads)
    cells = forward.mesh.cell_count
    ids = np.full(cells, -1)
    ids[: cells // 2] = np.arange(cells // 2) // 3
    parameters = int(ids.max()) + 2  # the last column has no cell
    expected = ref.jacobian(sigma, ids=ids, count=parameters)
    _, aggregated = forward.solve_with_jacobian(torch.as_tensor(sigma), jacobian_cell_parameter_ids=ids,
                                                jacobian_parameter_count=parameters)
    print(f"[{dim}] aggregated jacobian vs oracle {relmax(aggregated, expected):.2e}")
    assert aggregated.shape == (forward.survey.measurement_count, parameters)
    assert relmax(aggregated, expected) < 1e-8 and np.all(aggregated.numpy()[:, -1] == 0)
    cot = np.random.default_rng(4).standard_normal(forward.survey.measurement_count)
    vjp = forward.normal_vjp(torch.as_tensor(sigma), torch.as_tensor(cot), cell_parameter_ids=ids,
                             parameter_count=parameters)
    assert relmax(vjp, expected.T @ cot) < 1e-8


def test_3d_relief_uses_numerical_geometric_factors_against_the_oracle():
    (nodes, tets, el, quads, s), forward = box_operator(relief=0.8, element_order=2)
    assert not forward.mesh.is_flat_surface and forward.geometric_factor_mode == "auto"
    ref = oracle3d.Forward3D(nodes, tets, el, quads, element_order=2)
    homogeneous = forward.apparent_resistivity_values(torch.full((forward.mesh.cell_count,), 0.01)).numpy()
    assert relmax(homogeneous, 100.0) < 1e-8
    rhoa = forward.apparent_resistivity_values(torch.as_tensor(s)).numpy()
    assert relmax(rhoa, ref.apparent_resistivity(s)) < 1e-6


def test_mixed_precision_sensitivity_stays_within_1e5_of_the_oracle():
    case, forward = flat_operator(sensitivity_precision="mixed")
    ref = oracle2p5d.Forward2p5D(case.nodes, case.cells, case.electrodes, case.measurements, case.surface_node_ids)
    jac = forward.jacobian(torch.as_tensor(case.conductivity)).numpy()
    err = relmax(jac, ref.jacobian(case.conductivity))
    print(f"mixed-precision jacobian max rel err {err:.2e}")
    assert err < 1e-5
    again = forward.jacobian(torch.as_tensor(case.conductivity)).numpy()
    assert np.array_equal(again, jac)  # bitwise repeatable


def test_validation_errors_and_close():
    case, forward = flat_operator()
    with pytest.raises(ValueError, match="topographic_geometric_factor_mode"):
        flat_operator(topographic_geometric_factor_mode="x")
    with pytest.raises(ValueError, match="non-negative"):
        flat_operator(normal_field_cache_max_entries=-1)
    with pytest.raises(ValueError, match="conductivity must be scalar"):
        forward.solve(torch.ones(3, dtype=torch.float64))
    with pytest.raises(ValueError, match="parameter_count requires"):
        forward.normal_vjp(torch.as_tensor(case.conductivity), torch.ones(forward.survey.measurement_count),
                           parameter_count=3)
    sigma = torch.as_tensor(case.conductivity)
    first = forward.resistance(sigma).numpy()
    forward.close()
    assert relmax(forward.resistance(sigma).numpy(), first) < 1e-12
    (_, _, el, q, _), _ = box_operator()
    mesh = adt.Mesh3D.from_arrays(*adt.cases.box_mesh())
    with pytest.raises(ValueError, match="boundary_mode"):
        adt.ERTForward3D.from_mesh_survey(mesh, adt.Survey.from_arrays(el, q), boundary_mode="x")
    with pytest.raises(ValueError, match="element_order"):
        adt.ERTForward3D.from_mesh_survey(mesh, adt.Survey.from_arrays(el, q), element_order=3)
    with pytest.raises(ValueError, match="three-coordinate"):
        adt.ERTForward3D.from_mesh_survey(mesh, adt.Survey.from_arrays(el[:, :2], q))


def test_terrain_cache_dir_stores_and_reuses_the_unit_primary(tmp_path, monkeypatch):
    """With ``terrain_cache_dir`` the host-computed unit primary goes to ``<sha256>.npz`` and a
    second operator on the same inputs reads it back instead of recomputing it."""

    case, forward = flat_operator(terrain_cache_dir=tmp_path)
    sigma = torch.as_tensor(case.conductivity)
    first = forward.apparent_resistivity_values(sigma).numpy()
    files = list(tmp_path.glob("*.npz"))
    assert len(files) == 1
    _, again = flat_operator(terrain_cache_dir=tmp_path)
    monkeypatch.setattr(again, "_analytic_primary", lambda: pytest.fail("the cached primary was recomputed"))
    assert np.array_equal(again.apparent_resistivity_values(sigma).numpy(), first)
    forward.close()
    again.close()
