"""Meshes of different sizes in one batched solve (``MixedMeshBatch``): each operator's apparent
resistivities against its own solve and the FP64 oracle, with the multigrid and the learned
preconditioner (the learned damping runs once per CSR pattern of the plan), series of two
models, and operators with different electrode counts (padded source columns)."""

from __future__ import annotations

import importlib

import numpy as np
import pytest
import torch

from oracle import ert2p5d as oracle2p5d
from tests.conftest import adt, relmax

pytestmark = pytest.mark.gpu
batch = importlib.import_module("ad-tlert_amd.batch")


@pytest.fixture(scope="module")
def meshes():
    cases = adt.cases.mixed_meshes_ws(node_targets=(2000, 4000, 7000), electrodes=24, rows=150)
    small = adt.cases.mixed_meshes_ws(seed=9, node_targets=(1500,), electrodes=16, rows=60)[0]
    return cases + [small]


def operators(cases, preconditioner):
    ops = []
    for c in cases:
        mesh = adt.Mesh.from_arrays(c.nodes, c.cells)
        survey = adt.Survey.from_arrays(c.electrodes, c.measurements)
        ops.append(adt.ERTForward2p5D.from_mesh_survey(mesh, survey, preconditioner=preconditioner))
    return ops


@pytest.mark.parametrize("preconditioner", ["amg", "learned"])
def test_mixed_batch_matches_single_solves_and_the_oracle(meshes, preconditioner):
    ops = operators(meshes, preconditioner)
    mixed = batch.MixedMeshBatch(ops)
    sigmas = [torch.as_tensor(c.conductivity) for c in meshes]
    together = mixed.apparent_resistivity_values(sigmas)
    stats = mixed.solver.last
    assert stats.preconditioner == preconditioner and float(stats.relative_residual.max()) <= 1e-10
    assert mixed.solver.last_plan.n_sys == sum(op._dev.W for op in ops)
    for k, (c, op) in enumerate(zip(meshes, ops)):
        alone = op.apparent_resistivity_values(sigmas[k]).numpy()
        ref = oracle2p5d.Forward2p5D(c.nodes, c.cells, c.electrodes, c.measurements).apparent_resistivity(
            c.conductivity)
        err_alone, err_oracle = relmax(together[k].numpy(), alone), relmax(together[k].numpy(), ref)
        print(f"[{preconditioner}] mesh {k}: {c.nodes.shape[0]} nodes, {op._dev.E} electrodes: vs own solve "
              f"{err_alone:.2e}, vs oracle {err_oracle:.2e}")
        assert err_alone < 1e-8 and err_oracle < 1e-6
    mixed.close()
    for op in ops:
        op.close()


def test_mixed_batch_of_series(meshes):
    ops = operators(meshes[:2], "amg")
    mixed = batch.MixedMeshBatch(ops)
    series = [torch.as_tensor(np.stack((c.conductivity, 1.7 * c.conductivity))) for c in meshes[:2]]
    out = mixed.apparent_resistivity_values(series)
    for op, s, r in zip(ops, series, out):
        assert r.shape == (2, op.survey.measurement_count)
        assert relmax(r[1].numpy(), r[0].numpy() / 1.7) < 1e-8  # rho_a scales with 1 / sigma
        assert relmax(r.numpy(), op.apparent_resistivity_series(s).numpy()) < 1e-8
    with pytest.raises(ValueError, match="same number of models"):
        mixed.resistance([series[0], series[1][0]])
    mixed.close()
