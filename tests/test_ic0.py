"""IC(0)-preconditioned batched CG.

CPU: the IC(0) oracle pinned (the sequential factor against the defining property (L L^T)_ij = a_ij
on the pattern and against Cholesky where IC(0) has no dropped fill; the fixed-point factor against
the sequential one; the sweep solves against exact triangular solves and for symmetry; PCG against
a dense solve) and the host-built lower/transposed patterns.
GPU: ``adt_ic0_pcg`` against the oracle -- the same first iterates (the factor, the sweeps and the
CG recurrences all enter them), convergence at 1e-10 with the oracle's iteration counts, a dense
solve, bitwise repeatability, fewer iterations than Jacobi -- and the 2.5D operator with
``preconditioner="ic0"`` against the FP64 oracle on the C1 strip."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import torch

from oracle import cpu_pcg, ic0 as oic0
from tests.conftest import adt, relmax

native = adt.native


def grid_matrix(nx, ny, seed, mass=0.05):
    """SPD 5-point operator with random positive edge conductances plus a mass term."""

    rng = np.random.default_rng(seed)
    n = nx * ny
    idx = np.arange(n).reshape(ny, nx)
    edges = np.concatenate((np.stack((idx[:, :-1].ravel(), idx[:, 1:].ravel()), 1),
                            np.stack((idx[:-1, :].ravel(), idx[1:, :].ravel()), 1)))
    w = np.exp(rng.uniform(-1.5, 1.5, len(edges)))
    i, j = edges[:, 0], edges[:, 1]
    lap = sp.coo_matrix((np.concatenate((-w, -w)), (np.concatenate((i, j)), np.concatenate((j, i)))), shape=(n, n))
    deg = -np.asarray(lap.sum(axis=1)).ravel()
    a = (lap + sp.diags(deg + mass * np.exp(rng.uniform(-1, 1, n)))).tocsr()
    a.sort_indices()
    return a


def test_exact_ic0_has_the_defining_property_and_is_cholesky_without_fill():
    a = grid_matrix(6, 5, 0)
    low = oic0.exact_ic0(a)
    prod = (low @ low.T).tocsr()
    lower = oic0.lower_pattern(a)
    rows = np.repeat(np.arange(a.shape[0]), np.diff(lower.indptr))
    got = np.asarray(prod[rows, lower.indices]).ravel()
    assert np.allclose(got, lower.data, rtol=1e-13, atol=1e-13)  # (L L^T)_ij = a_ij on the pattern
    assert not np.allclose((prod - a).toarray(), 0.0)  # fill was dropped: not the exact factor
    tri = sp.diags([np.full(9, -1.0), np.full(10, 2.5), np.full(9, -1.0)], [-1, 0, 1], format="csr")
    assert np.allclose(oic0.exact_ic0(tri).toarray(), np.linalg.cholesky(tri.toarray()), rtol=0, atol=1e-14)


def test_fixed_point_factor_reaches_the_sequential_factor():
    a = grid_matrix(5, 4, 1)
    exact = oic0.exact_ic0(a)
    few = oic0.chow_patel_factor(a, 3)
    full = oic0.chow_patel_factor(a, a.shape[0])
    assert relmax(full.data, exact.data) < 1e-13
    err3 = np.abs(few.data - exact.data).max()
    err1 = np.abs(oic0.chow_patel_factor(a, 1).data - exact.data).max()
    assert err3 < err1 and err3 < 0.2 * np.abs(exact.data).max()  # the sweeps converge


def test_sweep_solves_are_triangular_solves_and_symmetric():
    a = grid_matrix(5, 4, 2)
    n = a.shape[0]
    factor = oic0.exact_ic0(a)
    rng = np.random.default_rng(3)
    r = rng.standard_normal((n, 3))
    dense = factor.toarray()
    y = sla.solve_triangular(dense, r, lower=True)
    z = sla.solve_triangular(dense.T, y, lower=False)
    assert relmax(oic0.apply(factor, r, n - 1), z) < 1e-12
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    for sweeps in (0, 1, 2):
        mu, mv = oic0.apply(factor, u, sweeps), oic0.apply(factor, v, sweeps)
        assert abs(v @ mu - u @ mv) < 1e-12 * abs(v @ mu)
        assert u @ mu > 0 and v @ mv > 0


def test_oracle_pcg_solves_and_beats_jacobi():
    a = grid_matrix(20, 15, 4)
    b = np.random.default_rng(5).standard_normal((a.shape[0], 2))
    x, iters = oic0.pcg(a, b, tol=1e-10)
    assert relmax(x, np.linalg.solve(a.toarray(), b)) < 1e-8
    _, jac = cpu_pcg.jacobi_pcg(a, b, tol=1e-10)
    assert np.all(iters < 0.7 * np.asarray(jac))


def test_host_patterns_are_the_lower_triangle_and_its_transpose():
    a = grid_matrix(7, 3, 6)
    pat = native.ic0_patterns(a.indptr, a.indices)
    low = oic0.lower_pattern(a)
    assert np.array_equal(pat["lptr"], low.indptr) and np.array_equal(pat["lidx"], low.indices)
    assert np.array_equal(a.data[pat["lsrc"]], low.data)
    up = low.T.tocsr()
    up.sort_indices()
    assert np.array_equal(pat["uptr"], up.indptr) and np.array_equal(pat["uidx"], up.indices)
    positions = sp.csr_matrix((np.arange(low.nnz, dtype=float) + 1, low.indices, low.indptr), shape=low.shape)
    t = positions.T.tocsr()
    t.sort_indices()
    assert np.array_equal(pat["upos"], t.data.astype(int) - 1)
    bad = sp.csr_matrix(np.array([[1.0, 2.0], [2.0, 0.0]]))
    bad.eliminate_zeros()
    with pytest.raises(ValueError, match="diagonal"):
        native.ic0_patterns(bad.indptr, bad.indices)


def device_batch(mats, rhs, r, rp, x0=None):
    """Device tensors of systems sharing one pattern: values (n_sys, nnz), interleaved b and x."""

    dev = torch.device("cuda")
    n = mats[0].shape[0]
    values = torch.as_tensor(np.stack([m.data for m in mats]), device=dev)
    b = np.zeros((len(rhs), n, rp))
    for k, blk in enumerate(rhs):
        b[k, :, :r] = blk
    x = np.zeros((len(mats), n, rp)) if x0 is None else x0.copy()
    pat = {k: (torch.as_tensor(v, device=dev) if isinstance(v, np.ndarray) else v)
           for k, v in native.ic0_patterns(mats[0].indptr, mats[0].indices).items()}
    return (torch.as_tensor(mats[0].indptr.astype(np.int32), device=dev),
            torch.as_tensor(mats[0].indices.astype(np.int32), device=dev), pat, values,
            torch.as_tensor(b.reshape(-1), device=dev), torch.as_tensor(x.reshape(-1), device=dev))


def run(mats, rhs, r, rp, tol, max_iter, x0=None, **k):
    indptr, indices, pat, values, b, x = device_batch(mats, rhs, r, rp, x0)
    rc, iters, relres = native.ic0_pcg(len(mats), mats[0].shape[0], r, rp, indptr, indices, pat, values, b, len(rhs),
                                       x, tol, max_iter, **k)
    torch.cuda.synchronize()
    return rc, iters, relres, x.cpu().numpy().reshape(len(mats), mats[0].shape[0], rp)


@pytest.fixture(scope="module")
def systems():
    base = grid_matrix(40, 30, 7)
    mats = []
    for s in range(3):  # same pattern, different coefficients
        m = grid_matrix(40, 30, 10 + s)
        assert np.array_equal(m.indices, base.indices) and np.array_equal(m.indptr, base.indptr)
        mats.append(m)
    rng = np.random.default_rng(8)
    rhs = [rng.standard_normal((base.shape[0], 5))]
    rhs[0][:, 3] = 0.0  # a zero column: nothing to solve
    return mats, rhs


@pytest.mark.gpu
def test_device_ic0_first_iterates_equal_the_oracle(systems):
    mats, rhs = systems
    for sweeps in ((3, 2), (1, 0), (4, 3)):
        rc, iters, _, x = run(mats, rhs, 5, 8, 1e-30, 5, factor_sweeps=sweeps[0], solve_sweeps=sweeps[1])
        assert rc == native.ADT_ERR_NOT_CONVERGED
        for s, m in enumerate(mats):
            xo, io = oic0.pcg(m, rhs[0], tol=1e-30, max_iter=5, factor_sweeps=sweeps[0], solve_sweeps=sweeps[1])
            print(f"[sweeps {sweeps}] system {s}: x vs oracle after 5 iterations {relmax(x[s, :, :5], xo):.2e}")
            assert relmax(x[s, :, :5], xo) < 1e-11 and np.array_equal(iters[s], io)
            assert np.all(x[s, :, 5:] == 0.0)


@pytest.mark.gpu
def test_device_ic0_converges_like_the_oracle_and_beats_jacobi(systems):
    mats, rhs = systems
    rc, iters, relres, x = run(mats, rhs, 5, 8, 1e-10, 5000)
    assert rc == native.ADT_OK and float(relres.max()) <= 1e-10
    rc2, _, _, x2 = run(mats, rhs, 5, 8, 1e-10, 5000)
    assert rc2 == native.ADT_OK and np.array_equal(x, x2)  # fixed-order sums: bitwise repeatable
    for s, m in enumerate(mats):
        exact = np.linalg.solve(m.toarray(), rhs[0])
        xo, io = oic0.pcg(m, rhs[0], tol=1e-10)
        _, jac = cpu_pcg.jacobi_pcg(m, rhs[0], tol=1e-10)
        res = np.linalg.norm(rhs[0] - m @ x[s, :, :5], axis=0)
        print(f"system {s}: iterations device {iters[s].tolist()} oracle {io.tolist()} Jacobi {list(jac)}, "
              f"x vs dense {relmax(x[s, :, :5], exact):.2e}")
        assert np.all(res <= 1.01e-10 * np.linalg.norm(rhs[0], axis=0))
        assert relmax(x[s, :, :5], exact) < 1e-7
        assert np.all(np.abs(iters[s] - io) <= 1)
        assert iters[s][3] == 0 and np.all(x[s, :, 3] == 0.0)
        live = [0, 1, 2, 4]
        assert np.all(iters[s][live] < 0.6 * np.asarray(jac)[live])


@pytest.mark.gpu
def test_device_ic0_ignores_the_workspace_contents(systems):
    """The first search direction is the preconditioned residual itself: a workspace left full
    of NaN by an earlier call does not leak into it (once it did, through 0 * NaN)."""

    mats, rhs = systems
    indptr, indices, pat, values, b, x = device_batch(mats, rhs, 5, 8)
    n = mats[0].shape[0]
    size = int(native.library().adt_ic0_work_doubles(len(mats), n, int(indices.numel()), int(pat["lidx"].numel()), 8))
    work = torch.full((size,), float("nan"), dtype=torch.float64, device=values.device)
    rc, iters, relres = native.ic0_pcg(len(mats), n, 5, 8, indptr, indices, pat, values, b, 1, x, 1e-10, 5000,
                                       work=work)
    clean = run(mats, rhs, 5, 8, 1e-10, 5000)
    assert rc == native.ADT_OK and float(relres.max()) <= 1e-10
    assert torch.equal(x.cpu(), torch.as_tensor(clean[3].reshape(-1))) and np.array_equal(iters, clean[1])


@pytest.mark.gpu
def test_device_ic0_own_right_hand_sides_and_warm_start(systems):
    mats, _ = systems
    rng = np.random.default_rng(11)
    n = mats[0].shape[0]
    rhs = [rng.standard_normal((n, 2)) for _ in mats]
    x0 = np.zeros((len(mats), n, 4))
    exact = [np.linalg.solve(m.toarray(), b) for m, b in zip(mats, rhs)]
    for s in range(len(mats)):
        x0[s, :, :2] = exact[s] * (1 + 1e-3 * rng.standard_normal(exact[s].shape))
    rc, iters, relres, x = run(mats, rhs, 2, 4, 1e-10, 5000, x0=x0)
    rc_cold, iters_cold, _, _ = run(mats, rhs, 2, 4, 1e-10, 5000)
    assert rc == native.ADT_OK and rc_cold == native.ADT_OK
    for s in range(len(mats)):
        assert relmax(x[s, :, :2], exact[s]) < 1e-7
    assert iters.sum() < iters_cold.sum()  # the warm start pays


@pytest.mark.gpu
def test_strip_operator_with_ic0_matches_the_oracle():
    from oracle import ert2p5d as oracle2p5d

    case = adt.cases.strip_dd_4k()
    mesh = adt.Mesh.from_arrays(case.nodes, case.cells)
    survey = adt.Survey.from_arrays(case.electrodes, case.measurements)
    fwd = adt.ERTForward2p5D.from_mesh_survey(mesh, survey, preconditioner="ic0")
    amg = adt.ERTForward2p5D.from_mesh_survey(mesh, survey)
    sigma = torch.as_tensor(case.conductivity)
    rho = fwd.apparent_resistivity_values(sigma).numpy()
    stats = fwd._solver.last
    assert stats.preconditioner == "ic0" and float(stats.relative_residual.max()) <= 1e-10
    ref = oracle2p5d.Forward2p5D(case.nodes, case.cells, case.electrodes, case.measurements).apparent_resistivity(
        case.conductivity)
    base = amg.apparent_resistivity_values(sigma).numpy()
    print(f"[C1 ic0] mean iterations {stats.iterations.mean():.1f}, vs oracle {relmax(rho, ref):.2e}, "
          f"vs AMG {relmax(rho, base):.2e}")
    assert relmax(rho, ref) < 1e-6 and relmax(rho, base) < 1e-8
    fwd.close()
    amg.close()


@pytest.mark.gpu
@pytest.mark.parametrize("order,boundary", [(1, "mixed"), (2, "dirichlet")])
def test_box_operator_with_ic0_matches_the_oracle(order, boundary):
    from oracle import ert3d as oracle3d

    case = adt.cases.box_case()
    mesh = adt.Mesh3D.from_arrays(case.nodes, case.cells)
    survey = adt.Survey.from_arrays(case.electrodes, case.measurements)
    fwd = adt.ERTForward3D.from_mesh_survey(mesh, survey, element_order=order, boundary_mode=boundary,
                                            preconditioner="ic0", solver_tolerance=1e-12)
    ref = oracle3d.Forward3D(case.nodes, case.cells, case.electrodes, case.measurements, boundary_mode=boundary,
                             element_order=order)
    rho = fwd.apparent_resistivity_values(torch.as_tensor(case.conductivity)).numpy()
    stats = fwd._solver.last
    err = relmax(rho, ref.apparent_resistivity(case.conductivity))
    print(f"[3d P{order} {boundary} ic0] mean iterations {stats.iterations.mean():.1f}, vs oracle {err:.2e}")
    assert stats.preconditioner == "ic0" and err < 1e-6
    fwd.close()
