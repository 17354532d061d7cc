"""Device CGLS (the Gauss-Newton linear step) against the FP64 CGLS oracle, least squares and the
damped normal equations; the Levenberg-Marquardt step on the C1 strip's log Jacobian stacked
over a first-order smoothness operator.  Mirrors ``inversion/optimizers.py:245-359``."""

from __future__ import annotations

import importlib

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import cgls as ocgls
from tests.conftest import adt, relmax

gn = importlib.import_module("ad-tlert_amd.gauss_newton")
par = importlib.import_module("ad-tlert_amd.parameterized")


def random_system(m, n, density, seed, dense_rows=0):
    rng = np.random.default_rng(seed)
    a = sp.random(m, n, density=density, random_state=seed, format="csr") + sp.eye(m, n, format="csr")
    if dense_rows:
        a = sp.vstack((sp.csr_matrix(rng.standard_normal((dense_rows, n))), a), format="csr")
    return a, rng.standard_normal(a.shape[0])


def test_linearized_step_validation_and_step_limit():
    assert np.array_equal(gn.limit_step(np.array([0.5, -2.0]), 1.0), np.array([0.25, -1.0]))
    assert np.array_equal(gn.limit_step(np.array([0.5, -0.2]), 1.0), np.array([0.5, -0.2]))
    cfg = gn.LinearizedConfig(linearized_solver="lsqr")
    with pytest.raises(NotImplementedError, match="gpu_cgls"):
        gn.solve_linearized(sp.eye(3, format="csr"), np.ones(3), cfg)
    with pytest.raises(NotImplementedError, match="linearized algorithms"):
        gn.linearized_step(sp.eye(3, format="csr"), np.ones(3), np.zeros(3), {},
                           gn.LinearizedConfig(optimization_algorithm="adam"))
    a, b = random_system(20, 8, 0.3, 1)
    assert np.allclose(gn.linearized_gradient(a, b), -(a.T @ b))


@pytest.mark.gpu
@pytest.mark.parametrize("dense_rows", [0, 48], ids=["wave-rows", "dense-rows"])
def test_device_cgls_matches_the_oracle_and_least_squares(dense_rows):
    a, b = random_system(2000 if dense_rows else 400, 1500 if dense_rows else 150, 0.02, 4, dense_rows)
    for tol in (1e-8, 1e-26):
        x, iters, status = gn.gpu_cgls(a, b, max_iterations=3000, tolerance=tol, return_info=True)
        xo, io = ocgls.cgls(a, b, 3000, tol)
        print(f"[{dense_rows} dense rows, tol {tol:g}] iterations device {iters} oracle {io}, "
              f"x vs oracle {relmax(x, xo):.2e}, {status}")
        # the contract at a tolerance: ||A^T (b - A x)||^2 < tol ||A^T b||^2 (recomputed here, in
        # FP64 on the host), reached within an iteration or so of the oracle (the reduction
        # orders differ, so the last iterate before the stop can too)
        s0, s = a.T @ b, a.T @ (b - a @ x)
        assert float(s @ s) < 2 * tol * float(s0 @ s0) and status == "converged"
        assert abs(iters - io) <= max(1, 0.01 * io)  # near the rounding floor the last iterations are noise
        again = gn.gpu_cgls(a, b, max_iterations=3000, tolerance=tol)
        assert np.array_equal(again, x)  # fixed-order sums: bitwise repeatable
    assert relmax(x, np.linalg.lstsq(a.toarray(), b, rcond=None)[0]) < 1e-9  # full column rank
    x0, i0, s0 = gn.gpu_cgls(a, np.zeros(a.shape[0]), max_iterations=10, tolerance=1e-8, return_info=True)
    assert i0 == 0 and s0 == "nothing to solve" and np.all(x0 == 0)
    xc, ic, sc = gn.gpu_cgls(a, b, max_iterations=3, tolerance=1e-30, return_info=True)
    assert ic == 3 and sc == "iteration limit" and relmax(xc, ocgls.cgls(a, b, 3, 1e-30)[0]) < 1e-10


@pytest.mark.gpu
def test_levenberg_marquardt_step_on_the_strip_jacobian():
    """One damped Gauss-Newton step of a log-resistivity inversion on C1 (3993 nodes, 1000 rows):
    A = [J / err; sqrt(lambda_s) R], b = [(d - f(m)) / err; -sqrt(lambda_s) R m] with R the
    first-order differences across cell interfaces; the device step against the oracle CGLS on
    the same stacked system (plus the LM rows) and the step limit."""

    case = adt.cases.strip_dd_4k()
    mesh = adt.Mesh.from_arrays(case.nodes, case.cells)
    survey = adt.Survey.from_arrays(case.electrodes, case.measurements)
    forward = adt.ERTForward2p5D.from_mesh_survey(mesh, survey)
    model = np.log(1.0 / case.conductivity)
    observed = np.log(forward.apparent_resistivity_values(torch.as_tensor(case.conductivity * 1.3)).numpy())
    start = np.full(model.size, float(np.mean(model)))
    predicted, jac = par.log_response_and_jacobian(forward, torch.as_tensor(np.exp(-start)))
    pairs, _ = par.neighbor_pairs(mesh)
    pairs = np.asarray(pairs)
    rows = np.repeat(np.arange(len(pairs)), 2)
    smooth = sp.csr_matrix((np.tile([1.0, -1.0], len(pairs)), (rows, pairs.reshape(-1))), shape=(len(pairs), model.size))
    err, lam = 0.03, 20.0
    a = sp.vstack((sp.csr_matrix(jac / err), np.sqrt(lam) * smooth), format="csr")
    b = np.concatenate(((observed - predicted) / err, -np.sqrt(lam) * (smooth @ start)))
    cfg = gn.LinearizedConfig(optimization_algorithm="levenberg_marquardt", lm_damping=1e-2, max_log_step=1.0,
                              cgls_max_iterations=400, cgls_tolerance=1e-8)
    stacked = sp.vstack((a, np.sqrt(cfg.lm_damping) * sp.eye(model.size, format="csr")), format="csr")
    rhs = np.concatenate((b, np.zeros(model.size)))
    # the same iterates as the oracle while rounding has not yet grown (this system is
    # ill-conditioned: over hundreds of iterations the two reduction orders drift apart)
    early = gn.gpu_cgls(stacked, rhs, max_iterations=10, tolerance=1e-30)
    assert relmax(early, ocgls.cgls(stacked, rhs, 10, 1e-30)[0]) < 1e-9
    step = gn.linearized_step(a, b, start, {}, cfg)
    raw, iters = ocgls.cgls(stacked, rhs, 400, 1e-8)
    expected = gn.limit_step(raw, 1.0)

    def objective(x):
        return float(np.linalg.norm(stacked @ x - rhs))

    device_raw = gn.gpu_cgls(stacked, rhs, max_iterations=400, tolerance=1e-8)
    print(f"[C1 LM step] {a.nnz} entries, oracle iterations {iters}, objective device {objective(device_raw):.9e} "
          f"oracle {objective(raw):.9e} start {np.linalg.norm(rhs):.6e}, step vs oracle {relmax(step, expected):.2e}")
    assert objective(device_raw) <= (1 + 1e-4) * objective(raw)
    s0, s_end = stacked.T @ rhs, stacked.T @ (rhs - stacked @ device_raw)
    assert float(s_end @ s_end) < 2e-8 * float(s0 @ s0)
    assert np.abs(step).max() <= 1.0 + 1e-12 and np.allclose(step, gn.limit_step(device_raw, 1.0))
    assert objective(device_raw) < 0.5 * np.linalg.norm(rhs)
    forward.close()
