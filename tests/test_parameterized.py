"""Parameterized consumers of the forward path: the parameter -> forward-cell maps (CPU), and
``ParameterizedERTForward2p5D.forward_and_jacobian`` / the matrix-free normal-vjp series on the
GPU against the FP64 oracle's per-cell Jacobian mapped the same way (fused column aggregation
under the prolongation rule, the projection otherwise).  Mirrors the reference's
``tests/test_inversion*.py`` use of ``inversion/parameterized.py`` and ``matrix_free.py``."""

from __future__ import annotations

import importlib

import numpy as np
import pytest
import torch

from oracle import ert2p5d as oracle2p5d
from tests.conftest import adt, relmax

par = importlib.import_module("ad-tlert_amd.parameterized")


def strip():
    """The quad strip with parameters on a central window, two neighbouring cells per parameter."""

    case = adt.cases.quad_strip_case()
    mesh = adt.Mesh.from_arrays(case.nodes, case.cells, surface_node_ids=case.surface_node_ids)
    centers = case.nodes[case.cells].mean(axis=1)
    window = np.flatnonzero((centers[:, 0] > 12) & (centers[:, 0] < 46) & (centers[:, 1] > -9))
    window = window[np.lexsort((centers[window, 0], centers[window, 1]))]
    ids = np.full(mesh.cell_count, -1, dtype=np.int32)
    ids[window] = np.arange(window.size) // 2
    reps = np.asarray([window[np.flatnonzero(ids[window] == p)[0]] for p in range(ids.max() + 1)])
    return case, mesh, ids, reps


def test_neighbor_pairs_known_answers():
    nodes, cells = adt.cases.grid_triangles(2, 1)  # 4 triangles: 3 interior edges
    mesh = adt.Mesh.from_arrays(nodes, cells)
    pairs, nz = par.neighbor_pairs(mesh)
    assert sorted(pairs) == [(0, 1), (1, 2), (2, 3)]
    assert all(a < b for a, b in pairs)
    # the vertical edge between the two squares has a horizontal normal; the diagonals 45 degrees
    by_pair = dict(zip(pairs, nz))
    assert by_pair[(1, 2)] == pytest.approx(0.0, abs=1e-15)
    assert by_pair[(0, 1)] == pytest.approx(np.sqrt(0.5)) and by_pair[(2, 3)] == pytest.approx(np.sqrt(0.5))
    box_nodes, tets = adt.cases.box_mesh(1, 1, 1)
    pairs3, nz3 = par.neighbor_pairs(adt.Mesh3D.from_arrays(box_nodes, tets))
    assert len(pairs3) == 6 and np.all((nz3 >= 0) & (nz3 <= 1))  # 6 tets around a diagonal share 6 faces


def host_only(ids, reps, mode):
    """The maps without a GPU: a stand-in operator carrying only the mesh and survey."""

    case = adt.cases.quad_strip_case()
    mesh = adt.Mesh.from_arrays(case.nodes, case.cells, surface_node_ids=case.surface_node_ids)
    fake = type("Op", (), {})()
    fake.mesh, fake.survey = mesh, adt.Survey.from_arrays(case.electrodes, case.measurements)
    return par.ParameterizedERTForward2p5D(fake, reps, forward_cell_parameter_ids=ids, background_mode=mode)


def test_parameter_maps_and_validation():
    _, mesh, ids, reps = strip()
    p = host_only(ids, reps, "pygimli_prolongation")
    P = p._resistivity_prolongation_matrix.toarray()
    active = ids >= 0
    assert np.array_equal(P[active], np.eye(p.cell_count)[ids[active]])
    assert np.allclose(P.sum(axis=1), 1.0) and P.min() >= 0.0  # background rows are weighted averages
    logs = np.random.default_rng(1).normal(3.0, 0.5, p.cell_count)
    full = p._full_log_model(logs)
    assert np.allclose(full[active], logs[ids[active]])
    assert np.all(full[~active] >= logs.min() - 1e-12) and np.all(full[~active] <= logs.max() + 1e-12)
    nearest = host_only(ids, reps, "nearest")
    full_n = nearest._full_log_model(logs)
    assert np.allclose(full_n[~active], logs[nearest._background_parameter_ids])
    mean = host_only(ids, reps, "fixed_mean")
    assert np.allclose(mean._full_log_model(logs)[~active], logs.mean())
    proj = mean._jacobian_projection.toarray()
    assert np.allclose(proj.sum(axis=1), 1.0)
    t = torch.as_tensor(np.stack((logs, logs + 0.1)))
    assert np.allclose(p.log_model_to_full(t)[0].numpy(), full)
    assert np.allclose(nearest.log_model_to_full(t)[1].numpy(), nearest._full_log_model(logs + 0.1))
    with pytest.raises(ValueError, match="background_mode"):
        host_only(ids, reps, "other")
    with pytest.raises(ValueError, match="ordered representatives"):
        host_only(ids, reps[::-1], "nearest")
    with pytest.raises(ValueError, match="must be unique"):
        host_only(ids, np.concatenate((reps[:1], reps[:1])), "nearest")
    with pytest.raises(ValueError, match="resistivity_model must have shape"):
        par.as_log_model(np.ones(3), 4, False, "resistivity_model")
    with pytest.raises(ValueError, match="must contain positive resistivity values"):
        par.as_log_model(-np.ones(4), 4, False, "resistivity_model")


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["pygimli_prolongation", "nearest", "fixed_mean"])
def test_forward_and_jacobian_and_matrix_free_vjp_match_the_oracle(mode):
    case, mesh, ids, reps = strip()
    survey = adt.Survey.from_arrays(case.electrodes, case.measurements)
    forward = par.ParameterizedERTForward2p5D.from_mesh_survey(mesh, survey, reps, forward_cell_parameter_ids=ids,
                                                               background_mode=mode)
    ref = oracle2p5d.Forward2p5D(case.nodes, case.cells, case.electrodes, case.measurements, case.surface_node_ids)
    logs = np.log(1.0 / case.conductivity[reps]) + np.random.default_rng(2).normal(0.0, 0.2, forward.cell_count)
    predicted, jac = forward.forward_and_jacobian(logs)
    full_log, projection = forward._full_log_model_and_projection(logs)
    sigma = np.exp(-full_log)
    rhoa = ref.apparent_resistivity(sigma)
    k = ref.geometric_factors()
    per_cell = -(k / rhoa)[:, None] * ref.jacobian(sigma) * sigma[None, :]  # d log rhoa / d log rho_c
    if mode == "pygimli_prolongation":  # fused: parameter columns sum their cells, background dropped
        onehot = np.zeros((ref.C, forward.cell_count))
        onehot[ids >= 0, ids[ids >= 0]] = 1.0
        expected = per_cell @ onehot
    else:
        expected = np.asarray((projection.T @ per_cell.T).T)
    print(f"[{mode}] log rhoa {relmax(predicted, np.log(rhoa)):.2e}, jacobian {relmax(jac, expected):.2e}")
    assert relmax(predicted, np.log(rhoa)) < 1e-7 and relmax(jac, expected) < 1e-5
    assert np.allclose(forward.forward(logs), predicted, rtol=0, atol=1e-12)
    g = np.random.default_rng(3).standard_normal((2, ref.D))
    rows = par.normal_log_response_vjp_series(forward, np.stack((logs, logs)), np.stack((predicted, predicted)), g)
    assert relmax(rows, g @ jac) < 1e-8
    forward.close()
