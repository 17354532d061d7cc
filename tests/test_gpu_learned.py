"""Learned multigrid damping on the GPU: network forward/backward against the FP64 oracle
(``oracle/gnn.py``), repeatability, the constant-damping identity, solves through the learned
preconditioner, and training on the GPU."""

import importlib

import numpy as np
import pytest
import torch

from oracle import gnn as ognn
from tests.conftest import adt, relmax
from tests.test_gpu_kernels import random_spd_batch

pytestmark = pytest.mark.gpu
native = adt.native
learned = importlib.import_module("ad-tlert_amd.learned")


def dev():
    return torch.device("cuda", 0)


def problem(side=12, seed=4):
    rng = np.random.default_rng(seed)
    indptr, indices, values, dense = random_spd_batch(rng, side=side, S=2)
    agg = native.amg_aggregates(indptr, indices, values[0])
    graph = learned.GnnGraph.from_pattern(indptr, indices, agg, dev())
    flags = graph.flags.cpu().numpy().astype(np.int64)
    return indptr, indices, values, dense, graph, flags


def random_params(seed=1):
    rng = np.random.default_rng(seed)
    flat = learned._he_init(seed).astype(np.float64)
    o = 0
    for name, shape in learned.LAYOUT:
        size = int(np.prod(shape))
        if name in ("v", "d", "c") or name.startswith("b"):
            flat[o:o + size] = rng.standard_normal(size) * 0.3
        o += size
    return flat.astype(np.float32)


def test_forward_matches_the_oracle():
    indptr, indices, values, _, graph, flags = problem()
    params = random_params()
    net = learned.DampingNetwork(params)
    omega_p = np.array([0.61, 0.7])
    out = net.omega(graph, torch.as_tensor(values, device=dev()), omega_p).cpu().numpy()
    for s in range(2):
        expected, _ = ognn.forward(params.astype(np.float64), indptr, indices, values[s], flags, omega_p[s])
        err = np.abs(out[s] - expected).max() / omega_p[s]
        print(f"system {s}: max |omega - oracle| / omega_p = {err:.2e}")
        assert err < 1e-4
    assert np.ptp(out[0]) > 1e-3  # the network is not trivially constant


def test_backward_matches_the_oracle_and_is_repeatable():
    indptr, indices, values, _, graph, flags = problem(seed=5)
    params = random_params(2)
    net = learned.DampingNetwork(params)
    omega_p = 0.65
    vals = torch.as_tensor(values[0], device=dev())
    rng = np.random.default_rng(3)
    domega = rng.standard_normal(values.shape[1])

    def grads():
        master = net.master.detach().clone().requires_grad_(True)
        omega = learned._Damping.apply(master, net, graph, vals, omega_p)
        omega.backward(torch.as_tensor(domega, device=dev()))
        return omega.detach(), master.grad.detach()

    omega, g = grads()
    _, cache = ognn.forward(params.astype(np.float64), indptr, indices, values[0], flags, omega_p)
    expected = ognn.backward(cache, domega)
    err = np.abs(g.cpu().numpy() - expected).max() / np.abs(expected).max()
    print(f"gradient max error / max |gradient| = {err:.2e}")
    assert err < 2e-3
    for _ in range(3):
        o2, g2 = grads()
        assert torch.equal(o2, omega) and torch.equal(g2, g)


def test_untrained_network_returns_the_constant_damping():
    _, _, values, _, graph, _ = problem()
    net = learned.DampingNetwork(seed=3)
    out = net.omega(graph, torch.as_tensor(values, device=dev()), [0.6, 0.6])
    assert torch.all(out == 0.6)


def test_learned_preconditioner_solves_to_tolerance():
    from tests.test_gpu_kernels import solve_batch

    _, stats = solve_batch(native.PRECOND_LEARNED, True, omega=0.5)  # per-slot damping array path
    indptr, indices, values, dense, graph, _ = problem(side=20, seed=2)
    net = learned.DampingNetwork(random_params(4))
    solver = adt.solver.BatchedSolver(preconditioner="learned", damping_network=net, amg_coarse_max=16)
    n, S, R = 400, 2, 13
    plan = solver.plan("t", np.full(S, n), R, pattern_offsets=(np.zeros(S, np.int64), np.zeros(S, np.int64),
                       np.arange(S, dtype=np.int64) * values.shape[1]), host_pattern=(indptr, indices))
    rng = np.random.default_rng(9)
    b = rng.standard_normal((S, n, plan.rp))
    b[:, :, R:] = 0.0
    x = torch.zeros((S, n, plan.rp), dtype=torch.float64, device=dev())
    stats = solver.solve(plan, torch.as_tensor(indptr, device=dev()), torch.as_tensor(indices, device=dev()),
                         torch.as_tensor(values, device=dev()), torch.as_tensor(b, device=dev()), x)
    assert stats.preconditioner == "learned" and stats.fallback_rhs == 0
    assert relmax(x.cpu().numpy()[:, :, :R], np.linalg.solve(dense, b[:, :, :R])) < 1e-7
    solver.close()


def test_training_on_the_gpu_reduces_the_two_grid_loss():
    indptr, indices, values, _, _, _ = problem(side=16, seed=6)
    rng = np.random.default_rng(0)
    systems = []
    for s in range(2):
        vals = torch.as_tensor(values[s], device=dev())
        errors = learned.error_samples(torch.as_tensor(indptr, device=dev()), torch.as_tensor(indices, device=dev()),
                                       vals, 256, 32, 15, rng, (indptr, indices))
        systems.append(learned.prepare_system(indptr, indices, vals, errors))
    net = learned.DampingNetwork(seed=0)
    before = [float(learned.two_grid_loss(s, net.omega(s.graph, s.values, s.omega_p))) for s in systems]
    losses = learned.train(net, systems, steps=40, lr=1e-2)
    after = [float(learned.two_grid_loss(s, net.omega(s.graph, s.values, s.omega_p))) for s in systems]
    print(f"two-grid loss {np.mean(before):.4f} -> {np.mean(after):.4f}")
    assert np.isfinite(losses).all()
    assert np.mean(after) < np.mean(before)


def test_trained_network_against_the_constant_damping():
    """What the committed, trained network changes against the constant damping on conductivity
    fields it never saw (seed 123; training used seed 7).  Measured: it lowers the two-grid
    energy ratio it was trained on by a few percent, and the full multigrid-preconditioned
    flexible CG needs the same number of iterations to within 2% (21.26 against 21.24 when the
    weights were committed): the cycle, not the network, makes the solver fast.  The test checks
    exactly that: a lower training objective and an iteration count equal within 2%."""

    net = learned.DampingNetwork.load()
    case = adt.cases.strip_dd_4k()
    mesh = adt.Mesh.from_arrays(case.nodes, case.cells)
    survey = adt.Survey.from_arrays(case.electrodes, case.measurements)
    op = adt.ERTForward2p5D.from_mesh_survey(mesh, survey, preconditioner="amg")
    op._solver.damping_network = net
    d = op._dev
    rng = np.random.default_rng(123)
    centers = case.nodes[case.cells].mean(axis=1)
    ratios, iters = [], {"amg": [], "learned": []}
    for f in range(4):
        sigma = adt.cases.log_conductivity_2d(centers, rng) if f % 2 == 0 else \
            10.0 ** (-(2.0 + 0.6 * adt.cases.smooth_fields_2d(centers, rng, 1)[0]))
        values = op._assemble_device(d, torch.as_tensor(sigma)[None])
        w = int(rng.integers(0, d.W))
        errors = learned.error_samples(d.indptr, d.indices, values[w], d.N, 32, 40, rng,
                                       (d.pattern.indptr, d.pattern.indices))
        system = learned.prepare_system(d.pattern.indptr, d.pattern.indices, values[w].contiguous(), errors)
        with torch.no_grad():
            constant = float(learned.two_grid_loss(system, torch.full_like(system.values, system.omega_p)))
            trained = float(learned.two_grid_loss(system, net.omega(system.graph, system.values, system.omega_p)))
        ratios.append(trained / constant)
        for name in ("amg", "learned"):
            op._solver.preconditioner = name
            op._solve_fields(torch.as_tensor(sigma)[None])
            assert op._solver.last.relative_residual.max() <= 1e-10
            iters[name].append(float(op._solver.last.iterations.mean()))
    print(f"two-grid energy ratio trained/constant: {np.round(ratios, 3).tolist()}; mean FCG iterations "
          f"constant {np.mean(iters['amg']):.2f}, learned {np.mean(iters['learned']):.2f}")
    assert np.mean(ratios) < 1.0
    assert abs(np.mean(iters["learned"]) - np.mean(iters["amg"])) <= 0.02 * np.mean(iters["amg"])
    op.close()
