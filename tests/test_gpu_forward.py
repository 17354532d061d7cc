"""2.5D and 3D operators on the GPU against the FP64 CPU oracle (same seeded inputs).

Bars: apparent resistivity within 1e-6 (max-norm relative), Jacobian / gradient entries within
1e-5, plus the reference suite's own invariants (``tests/test_forward_gpu.py``,
``tests/test_forward3d_gpu.py``): homogeneous half-space, scaling, dot tests, finite
differences, Jacobian vs matrix-free products, chunking and parameter aggregation.
"""

import numpy as np
import pytest
import torch

from oracle import ert2p5d as oracle2p5d
from oracle import ert3d as oracle3d
from tests.conftest import adt, relmax

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def quad():
    case = adt.cases.quad_strip_case()
    mesh = adt.Mesh.from_arrays(case.nodes, case.cells, surface_node_ids=case.surface_node_ids)
    survey = adt.Survey.from_arrays(case.electrodes, case.measurements)
    forward = adt.ERTForward2p5D.from_mesh_survey(mesh, survey, solver_tolerance=1e-12)
    ref = oracle2p5d.Forward2p5D(case.nodes, case.cells, case.electrodes, case.measurements, case.surface_node_ids)
    return case, forward, ref


@pytest.fixture(scope="module")
def tri():
    nodes, cells = adt.cases.grid_triangles(24, 10)
    electrodes = np.column_stack((np.arange(2.0, 23.0), np.zeros(21)))
    quads = adt.cases.dipole_dipole(21, spacings=(1, 2), limit=60)
    rng = np.random.default_rng(11)
    sigma = 10.0 ** rng.uniform(-2.5, -1.0, cells.shape[0])
    mesh = adt.Mesh.from_arrays(nodes, cells)
    survey = adt.Survey.from_arrays(electrodes, quads)
    forward = adt.ERTForward2p5D.from_mesh_survey(mesh, survey, solver_tolerance=1e-12)
    ref = oracle2p5d.Forward2p5D(nodes, cells, electrodes, quads)
    return sigma, forward, ref


def test_assembly_matches_the_oracle_and_is_bitwise_repeatable(quad):
    """Device assembly (Hilbert DOF numbering) equals the oracle's scatter-add, permuted."""

    import scipy.sparse as sp

    case, forward, ref = quad
    dev = forward._dev
    sigma = torch.as_tensor(case.conductivity)[None]
    values = forward._assemble_device(dev, sigma)
    perm = dev.perm_host
    for w in range(ref.W):
        a_dev = sp.csr_matrix((values[w].cpu().numpy(), dev.pattern.indices, dev.pattern.indptr),
                              shape=(ref.N, ref.N)).toarray()
        a_ref = ref.matrix(case.conductivity, w).toarray()[np.ix_(perm, perm)]
        assert relmax(a_dev, a_ref) < 1e-13
    for _ in range(5):
        assert torch.equal(forward._assemble_device(dev, sigma), values)


@pytest.mark.parametrize("name", ["quad", "tri"])
def test_apparent_resistivity_matches_the_oracle(name, quad, tri):
    if name == "quad":
        case, forward, ref = quad
        sigma = case.conductivity
    else:
        sigma, forward, ref = tri
    rhoa = forward.apparent_resistivity_values(torch.as_tensor(sigma)).numpy()
    expected = ref.apparent_resistivity(sigma)
    print(f"[{name}] rhoa max rel err {relmax(rhoa, expected):.3e}")
    assert relmax(rhoa, expected) < 1e-6


@pytest.mark.parametrize("name", ["quad", "tri"])
def test_multigrid_and_jacobi_solves_agree_and_multigrid_iterates_less(name, quad, tri):
    if name == "quad":
        case, forward, _ = quad
        sigma, mesh, survey = case.conductivity, forward.mesh, forward.survey
    else:
        sigma, forward, _ = tri
        mesh, survey = forward.mesh, forward.survey
    jacobi = adt.ERTForward2p5D.from_mesh_survey(mesh, survey, solver_tolerance=1e-12, preconditioner="jacobi")
    sig = torch.as_tensor(sigma) * 1.1  # not a cached field
    rhoa_j = jacobi.apparent_resistivity_values(sig).numpy()
    stats_j = jacobi._solver.last
    rhoa_a = forward.apparent_resistivity_values(sig).numpy()
    stats_a = forward._solver.last
    print(f"[{name}] iterations jacobi max {stats_j.iterations.max()} amg max {stats_a.iterations.max()}, "
          f"fallback {stats_a.fallback_rhs}")
    assert stats_j.preconditioner == "jacobi" and stats_a.preconditioner == "amg"
    assert relmax(rhoa_a, rhoa_j) < 1e-9
    assert stats_a.relative_residual.max() <= 1e-12 and stats_j.relative_residual.max() <= 1e-12
    assert stats_a.iterations.max() < stats_j.iterations.max()


def test_homogeneous_half_space_and_scaling(quad):
    case, forward, _ = quad
    homogeneous = forward.apparent_resistivity_values(torch.full((forward.mesh.cell_count,), 0.01)).numpy()
    assert relmax(homogeneous, 100.0) < 2e-3
    sigma = torch.as_tensor(case.conductivity)
    base = forward.apparent_resistivity_values(sigma).numpy()
    assert relmax(forward.apparent_resistivity_values(sigma / 3.0).numpy(), 3.0 * base) < 1e-9
    assert np.allclose(forward.apparent_resistivity_values(sigma, currents=2.0).numpy(), base / 2.0)


@pytest.mark.parametrize("name", ["quad", "tri"])
def test_normal_jacobian_matches_the_oracle(name, quad, tri, monkeypatch):
    if name == "quad":
        case, forward, ref = quad
        sigma = case.conductivity
    else:
        sigma, forward, ref = tri
    jac = forward.jacobian(torch.as_tensor(sigma)).numpy()
    expected = ref.jacobian(sigma)
    print(f"[{name}] jacobian max rel err {relmax(jac, expected):.3e}")
    assert relmax(jac, expected) < 1e-5
    # an explicit batch size launches the Gram kernel once per chunk of rows (the per-cell Gram
    # is recomputed per launch in the same order, so the rows are bitwise equal)
    launches = []
    gram = adt.native.sensitivity_gram
    monkeypatch.setattr(adt.native, "sensitivity_gram",
                        lambda *a, **k: launches.append(a[4].numel()) or gram(*a, **k))
    chunked = forward.jacobian(torch.as_tensor(sigma), batch_size=7).numpy()
    assert launches == [7] * (ref.D // 7) + ([ref.D % 7] if ref.D % 7 else [])
    assert np.array_equal(chunked, jac)


def test_vjp_jvp_and_normal_products_match_the_oracle(quad):
    case, forward, ref = quad
    sigma = torch.as_tensor(case.conductivity)
    rng = np.random.default_rng(3)
    g = rng.standard_normal(ref.D)
    d = rng.standard_normal(ref.C) * 1e-3
    assert relmax(forward.vjp(sigma, torch.as_tensor(g)).numpy(), ref.vjp(case.conductivity, g)) < 1e-5
    assert relmax(forward.jvp(sigma, torch.as_tensor(d)).numpy(), ref.jvp(case.conductivity, d)) < 1e-5
    assert relmax(forward.normal_vjp(sigma, torch.as_tensor(g)).numpy(), ref.normal_vjp(case.conductivity, g)) < 1e-5
    assert relmax(forward.normal_jvp(sigma, torch.as_tensor(d)).numpy(), ref.normal_jvp(case.conductivity, d)) < 1e-5
    lhs = float((forward.jvp(sigma, torch.as_tensor(d)).numpy() * g).sum())
    rhs = float((d * forward.vjp(sigma, torch.as_tensor(g)).numpy()).sum())
    assert lhs == pytest.approx(rhs, rel=1e-6)


def test_exact_jacobian_rows_match_the_oracle_vjp(quad):
    case, forward, ref = quad
    sigma = torch.as_tensor(case.conductivity)
    exact = forward.jacobian(sigma, include_robin_boundary_derivative=True, normal_sensitivity=False).numpy()
    g = np.random.default_rng(4).standard_normal(ref.D)
    assert relmax(exact.T @ g, ref.vjp(case.conductivity, g)) < 1e-5
    # scaling every resistivity scales the data: the exact log-derivative sums to one per row
    rhoa = forward.apparent_resistivity_values(sigma).numpy()
    resistance = forward.resistance(sigma).numpy()
    log_row_sums = -(exact * case.conductivity[None, :]).sum(axis=1) / resistance
    assert relmax(log_row_sums, np.ones_like(rhoa)) < 1e-6


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("boundary", ["mixed", "dirichlet"])
def test_3d_matches_the_oracle(order, boundary):
    case = adt.cases.box_case()
    mesh = adt.Mesh3D.from_arrays(case.nodes, case.cells)
    survey = adt.Survey.from_arrays(case.electrodes, case.measurements)
    forward = adt.ERTForward3D.from_mesh_survey(mesh, survey, element_order=order, boundary_mode=boundary,
                                                solver_tolerance=1e-12)
    ref = oracle3d.Forward3D(case.nodes, case.cells, case.electrodes, case.measurements, boundary_mode=boundary,
                             element_order=order)
    homogeneous = forward.apparent_resistivity_values(torch.full((mesh.cell_count,), 0.01)).numpy()
    assert relmax(homogeneous, 100.0) < 1e-6
    sigma = case.conductivity
    rhoa = forward.apparent_resistivity_values(torch.as_tensor(sigma)).numpy()
    print(f"[3d P{order} {boundary}] rhoa err {relmax(rhoa, ref.apparent_resistivity(sigma)):.3e}")
    assert relmax(rhoa, ref.apparent_resistivity(sigma)) < 1e-6
    jac = forward.jacobian(torch.as_tensor(sigma), include_robin_boundary_derivative=True).numpy()
    assert relmax(jac, ref.jacobian(sigma, robin=True)) < 1e-5
    chunked = forward.jacobian(torch.as_tensor(sigma), include_robin_boundary_derivative=True, batch_size=2).numpy()
    assert np.array_equal(chunked, jac)  # one launch per pair of rows, same per-row arithmetic
    rng = np.random.default_rng(order * 10 + (boundary == "dirichlet"))
    g, d = rng.standard_normal(ref.D), rng.standard_normal(ref.C) * 1e-3
    assert relmax(forward.vjp(torch.as_tensor(sigma), torch.as_tensor(g)).numpy(), ref.vjp(sigma, g)) < 1e-5
    assert relmax(forward.jvp(torch.as_tensor(sigma), torch.as_tensor(d)).numpy(), ref.jvp(sigma, d)) < 1e-5
    assert relmax(forward.normal_vjp(torch.as_tensor(sigma), torch.as_tensor(g)).numpy(), ref.normal_vjp(sigma, g)) < 1e-5
    assert relmax(forward.normal_jvp(torch.as_tensor(sigma), torch.as_tensor(d)).numpy(), ref.normal_jvp(sigma, d)) < 1e-5
