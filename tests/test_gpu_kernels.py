"""Native gfx950 kernels through the C ABI against plain references (GPU).

Mirrors the reference kernel suite (``tests/test_kernels_gpu.py:15-270``): the sensitivity
kernel against the einsum formula for K in {3, 4, 6, 10} at FP64 (1e-12) and FP32-product
(1e-4) precision, bitwise repeatability, the current stream, ragged tiles and empty
measurement sets; assembly and the multi-RHS product against host scatters and dense products.
"""

import numpy as np
import pytest
import torch

from tests.conftest import adt, relmax

pytestmark = pytest.mark.gpu
native = adt.native


def dev():
    return torch.device("cuda", 0)


def random_problem(k, W=5, S=9, N=60, C=40, D=17, seed=0):
    gen = torch.Generator().manual_seed(seed)
    draw = lambda *shape: torch.randn(*shape, generator=gen, dtype=torch.float64)  # noqa: E731
    phi = draw(W, S, N)
    stiff, mass = draw(C, k, k), draw(C, k, k)
    ws, wm = draw(W).abs(), draw(W).abs()
    cells = torch.randint(0, N, (C, k), generator=gen)
    a, b, m, n = (torch.randint(0, S + 1, (D,), generator=gen) for _ in range(4))  # S = remote
    return phi, a, b, m, n, cells, stiff, mass, ws, wm


def semidefinite_templates(C, k, seed):
    """Element-like templates: stiffness symmetric positive semidefinite with the constants in its
    null space (as every FEM stiffness), mass symmetric positive definite."""

    gen = torch.Generator().manual_seed(seed)
    a = torch.randn(C, k, k, generator=gen, dtype=torch.float64)
    centre = torch.eye(k, dtype=torch.float64) - 1.0 / k
    stiff = centre @ (a @ a.transpose(1, 2)) @ centre / k
    bm = torch.randn(C, k, k, generator=gen, dtype=torch.float64)
    mass = bm @ bm.transpose(1, 2) / k + 0.1 * torch.eye(k, dtype=torch.float64)
    return 0.5 * (stiff + stiff.transpose(1, 2)), 0.5 * (mass + mass.transpose(1, 2))


def einsum_reference(phi, a, b, m, n, cells, stiff, mass, ws, wm):
    padded = torch.cat((phi, torch.zeros_like(phi[:, :1])), dim=1)
    cur = (padded[:, a] - padded[:, b])[..., cells]
    rec = (padded[:, m] - padded[:, n])[..., cells]
    T = ws[:, None, None, None] * stiff[None] + wm[:, None, None, None] * mass[None]
    return -torch.einsum("wdci,wcij,wdcj->dc", rec, T, cur)


def run_sensitivity(problem, precision=0):
    phi, a, b, m, n, cells, stiff, mass, ws, wm = problem
    g = lambda t, dt=None: t.to(dev(), dtype=dt or t.dtype).contiguous()  # noqa: E731
    out = torch.empty((a.numel(), cells.shape[0]), dtype=torch.float64, device=dev())
    native.sensitivity(g(phi), *(g(x, torch.int32) for x in (a, b, m, n)), g(cells, torch.int32), g(stiff), g(mass),
                       ws.numpy(), wm.numpy(), out, precision)
    return out


@pytest.mark.parametrize("k", [3, 4, 6, 10])
@pytest.mark.parametrize("precision, tol", [(0, 1e-12), (1, 1e-4)])
def test_sensitivity_kernel_matches_the_einsum_reference(k, precision, tol):
    problem = random_problem(k)
    expected = einsum_reference(*problem)
    actual = run_sensitivity(problem, precision).cpu()
    assert actual.shape == expected.shape
    assert relmax(actual, expected) < tol


def test_sensitivity_kernel_is_repeatable_follows_the_stream_and_handles_empty_sets():
    problem = random_problem(10, C=301, D=33)
    first = run_sensitivity(problem)
    assert all(torch.equal(first, run_sensitivity(problem)) for _ in range(5))
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        on_stream = run_sensitivity(problem)
    stream.synchronize()
    assert torch.equal(on_stream, first)
    phi, a, b, m, n, cells, stiff, mass, ws, wm = problem
    empty = run_sensitivity((phi, a[:0], b[:0], m[:0], n[:0], cells, stiff, mass, ws, wm))
    assert empty.shape == (0, cells.shape[0])


def test_sensitivity_kernel_handles_ragged_tiles():
    problem = random_problem(3, C=37, D=5)
    assert relmax(run_sensitivity(problem).cpu(), einsum_reference(*problem)) < 1e-12


def test_multi_rhs_product_matches_dense_products_for_every_lane_width():
    rng = np.random.default_rng(1)
    n, nnz_row = 300, 7
    rows = np.repeat(np.arange(n), nnz_row)
    cols = rng.integers(0, n, rows.size)
    keys = np.unique(rows * n + cols)
    indptr = np.concatenate(([0], np.cumsum(np.bincount(keys // n, minlength=n)))).astype(np.int32)
    indices = (keys % n).astype(np.int32)
    S = 3
    values = rng.standard_normal((S, keys.size))
    for lanes, rp in ((1, 5), (8, 8), (16, 48), (32, 96), (64, 128)):
        x = rng.standard_normal((S, n, rp))
        y = torch.zeros((S, n, rp), dtype=torch.float64, device=dev())
        layout = (np.full(S, n, np.int32), np.zeros(S, np.int64), np.zeros(S, np.int64),
                  np.arange(S, dtype=np.int64) * keys.size, np.arange(S, dtype=np.int64) * n * rp, lanes, rp)
        native.spmm(layout, torch.as_tensor(indptr, device=dev()), torch.as_tensor(indices, device=dev()),
                    torch.as_tensor(values, device=dev()), torch.as_tensor(x, device=dev()), y, 0)
        dense = np.zeros((S, n, n))
        r = keys // n
        for s in range(S):
            dense[s, r, indices] = values[s]
        assert relmax(y.cpu().numpy(), dense @ x) < 1e-13, lanes


def random_spd_batch(rng, side=20, S=2):
    """2D Laplacian-like SPD matrices with random positive coefficients: CSR pattern, values
    (S, nnz) and dense copies."""

    n = side * side
    idx = np.arange(n).reshape(side, side)
    rows, cols = [], []
    for di, dj in ((0, 0), (0, 1), (1, 0), (0, -1), (-1, 0)):
        i0, j0 = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
        i1, j1 = i0 + di, j0 + dj
        ok = (i1 >= 0) & (i1 < side) & (j1 >= 0) & (j1 < side)
        rows.append(idx[i0[ok], j0[ok]])
        cols.append(idx[i1[ok], j1[ok]])
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    keys = np.unique(rows * n + cols)
    r, c = keys // n, keys % n
    indptr = np.concatenate(([0], np.cumsum(np.bincount(r, minlength=n)))).astype(np.int32)
    values = np.zeros((S, keys.size))
    dense = np.zeros((S, n, n))
    for s in range(S):
        w = rng.uniform(0.5, 2.0, keys.size)
        off = r != c
        sym = {}
        for k in np.flatnonzero(off):
            a_, b_ = min(r[k], c[k]), max(r[k], c[k])
            sym.setdefault((a_, b_), -w[k])
        for k in np.flatnonzero(off):
            values[s, k] = sym[(min(r[k], c[k]), max(r[k], c[k]))]
        dense[s, r[off], c[off]] = values[s, off]
        diag = -dense[s].sum(axis=1) + 0.01
        values[s, ~off] = diag[r[~off]]
        dense[s, r[~off], c[~off]] = values[s, ~off]
    return indptr, c.astype(np.int32), values, dense


def solve_batch(precond, staged, stall_limit=64, side=20, lanes=16, R=13, seed=2, omega=None, rows_per_tile=32,
                coarse_max=16, smoother_scale=native.SMOOTHER_SCALE):
    rng = np.random.default_rng(seed)
    indptr, indices, values, dense = random_spd_batch(rng, side)
    S, n = values.shape[0], side * side
    plan = native.PcgPlan(np.full(S, n), R, lanes, rows_per_tile, np.zeros(S), np.zeros(S), np.arange(S) * values.shape[1],
                          host_pattern=(indptr, indices) if staged or precond != native.PRECOND_JACOBI else None,
                          lds_bytes=8192 if staged else 0)
    assert (plan.lds_rows > 0) == staged
    values_d = torch.as_tensor(values, device=dev())
    data = None
    if precond != native.PRECOND_JACOBI:
        assert plan.attach_amg(values_d, coarse_max=coarse_max, stall_limit=stall_limit,
                               smoother_scale=smoother_scale) >= (3 if coarse_max < 64 else 2)
        assert plan.amg_level(1)["rows"] < n
    if precond == native.PRECOND_LEARNED:
        data = torch.full_like(values_d, omega)
    b = rng.standard_normal((S, n, plan.rp))
    b[:, :, R:] = 0.0
    b[1, :, 3] = 0.0  # a zero right-hand side -> zero solution
    x = torch.zeros((S, n, plan.rp), dtype=torch.float64, device=dev())
    x[0, :, 5] = 1.0
    rc, iters, relres = plan.solve(torch.as_tensor(indptr, device=dev()), torch.as_tensor(indices, device=dev()),
                                   values_d, torch.as_tensor(b, device=dev()), x, 1e-10, 5000, precond=precond,
                                   precond_data=data)
    assert rc == 0, native.library().adt_last_error()
    assert relres.max() <= 1e-10  # true residuals (restarted once where the recurrence drifted)
    assert iters[1, 3] == 0 and float(x[1, :, 3].abs().max()) == 0.0
    sol = np.linalg.solve(dense, b[:, :, :R])
    assert relmax(x.cpu().numpy()[:, :, :R], sol) < 1e-7
    stats = plan.stats()
    plan.close()
    return iters, stats


def test_smoother_damping_scale_is_bounded_and_every_scale_solves():
    """The cycle is a symmetric positive definite preconditioner only while the Jacobi damping
    times lambda_max stays below 2: the native setup rejects scales outside (0, 1.5).  Inside the
    range the flexible CG solves to the tolerance (against dense solves) at the stand-alone
    smoother's damping (1.0) and at the library default."""

    rng = np.random.default_rng(3)
    indptr, indices, values, _ = random_spd_batch(rng, 12)
    S, n = values.shape[0], 144
    plan = native.PcgPlan(np.full(S, n), 8, 8, 32, np.zeros(S), np.zeros(S), np.arange(S) * values.shape[1],
                          host_pattern=(indptr, indices), lds_bytes=8192)
    for bad in (0.0, 1.5, -1.0):
        with pytest.raises(ValueError, match="smoother_scale"):
            plan.attach_amg(torch.as_tensor(values, device=dev()), coarse_max=16, smoother_scale=bad)
    plan.close()
    counts = {}
    for scale in (1.0, native.SMOOTHER_SCALE):
        iters, stats = solve_batch(native.PRECOND_AMG, True, smoother_scale=scale)
        assert stats["fallback_rhs"] == 0
        counts[scale] = float(iters[iters > 0].mean())
    print(f"mean FCG iterations by smoother scale: {counts}")


def test_prolongation_damping_is_four_thirds_over_the_largest_eigenvalue():
    """The finest prolongator is smoothed with omega_p = 4 / (3 lambda_max(D_F^-1 A_F)), A_F the
    filtered matrix (strong off-diagonals, weak ones lumped into the diagonal D_F).  The setup's
    Lanczos estimate of lambda_max matches a dense symmetric eigensolver."""

    rng = np.random.default_rng(8)
    indptr, indices, values, dense = random_spd_batch(rng, 14, S=1)
    n, theta = 196, 0.08
    plan = native.PcgPlan(np.array([n]), 8, 8, 32, np.zeros(1), np.zeros(1), np.zeros(1),
                          host_pattern=(indptr, indices), lds_bytes=8192)
    plan.attach_amg(torch.as_tensor(values, device=dev()), coarse_max=16, theta=theta)
    omega_p = plan.amg_level(0)["omega"]
    plan.close()
    A = dense[0]
    d = np.diag(A).copy()
    off = A - np.diag(d)
    strong = (off != 0) & (np.abs(off) >= theta * np.sqrt(np.outer(d, d)))
    df = d + np.where(strong, 0.0, off).sum(axis=1)
    df = np.where(df > 0, df, d)
    AF = np.where(strong, off, 0.0) + np.diag(df)
    s = 1.0 / np.sqrt(df)
    lam = np.linalg.eigvalsh(s[:, None] * AF * s[None, :]).max()
    print(f"omega_p {omega_p:.12f} against 4/(3 lambda_max) {4 / (3 * lam):.12f}")
    assert abs(omega_p - 4.0 / (3.0 * lam)) <= 1e-6 * omega_p


def test_given_finest_prolongator_is_validated_and_solves():
    """A given finest-level prolongator (here: each fine row of a 2D grid interpolated from the
    even-even grid points around it, bilinearly) replaces the aggregation on level 0; coarser
    levels aggregate the Galerkin operator.  The cycle solves to the tolerance against dense
    solves; prolongators with entries out of range, non-finite values or plans of several CSR
    patterns are rejected."""

    rng = np.random.default_rng(12)
    side = 20
    indptr, indices, values, dense = random_spd_batch(rng, side, S=2)
    n = side * side
    cs = side // 2
    rows, cols, vals = [], [], []
    for i in range(side):
        for j in range(side):
            wi = [(i // 2, 1.0)] if i % 2 == 0 else [(i // 2, 0.5)] + ([(i // 2 + 1, 0.5)] if i // 2 + 1 < cs else [])
            wj = [(j // 2, 1.0)] if j % 2 == 0 else [(j // 2, 0.5)] + ([(j // 2 + 1, 0.5)] if j // 2 + 1 < cs else [])
            for a, x in wi:
                for b, y in wj:
                    rows.append(i * side + j)
                    cols.append(a * cs + b)
                    vals.append(x * y)
    rows, cols, vals = map(np.asarray, (rows, cols, vals))
    order = np.lexsort((cols, rows))
    p_ptr = np.concatenate(([0], np.cumsum(np.bincount(rows, minlength=n)))).astype(np.int32)
    p_idx, p_val = cols[order].astype(np.int32), vals[order]
    R, S = 13, 2
    values_d = torch.as_tensor(values, device=dev())

    plan = native.PcgPlan(np.full(S, n), R, 16, 32, np.zeros(S), np.zeros(S), np.arange(S) * values.shape[1],
                          host_pattern=(indptr, indices), lds_bytes=8192)
    bad_idx = p_idx.copy()
    bad_idx[3] = cs * cs
    bad_val = p_val.copy()
    bad_val[5] = np.nan
    for bad in ((p_ptr, bad_idx, p_val, cs * cs), (p_ptr, p_idx, bad_val, cs * cs)):
        with pytest.raises(ValueError, match="prolongator"):
            plan.attach_amg(values_d, coarse_max=16, prolongator=bad)
    levels = plan.attach_amg(values_d, coarse_max=16, prolongator=(p_ptr, p_idx, p_val, cs * cs))
    assert levels >= 2 and plan.amg_level(1)["rows"] == cs * cs
    b = rng.standard_normal((S, n, plan.rp))
    b[:, :, R:] = 0.0
    x = torch.zeros((S, n, plan.rp), dtype=torch.float64, device=dev())
    rc, iters, relres = plan.solve(torch.as_tensor(indptr, device=dev()), torch.as_tensor(indices, device=dev()),
                                   values_d, torch.as_tensor(b, device=dev()), x, 1e-10, 2000,
                                   precond=native.PRECOND_AMG)
    assert rc == 0 and relres.max() <= 1e-10 and plan.stats()["fallback_rhs"] == 0
    assert relmax(x.cpu().numpy()[:, :, :R], np.linalg.solve(dense, b[:, :, :R])) < 1e-7
    print(f"mean FCG iterations with the bilinear first level: {iters[:, :R].mean():.1f}")
    plan.close()
    pb, ib, vb, _ = random_spd_batch(rng, 13, S=1)
    two = native.PcgPlan(np.array([n, 169]), R, 16, 32, np.array([0, indptr.size]), np.array([0, indices.size]),
                         np.array([0, values.shape[1]]), host_pattern=(np.concatenate([indptr, pb]),
                                                                       np.concatenate([indices, ib])), lds_bytes=8192)
    with pytest.raises(ValueError, match="one CSR pattern"):
        two.attach_amg(torch.as_tensor(np.concatenate([values[0], vb[0]]), device=dev()), coarse_max=16,
                       prolongator=(p_ptr, p_idx, p_val, cs * cs))
    two.close()


def stencil_27_pattern(side):
    """CSR pattern of the 27-point stencil on a side^3 grid (rows of up to 27 entries)."""

    n = side ** 3
    g = np.arange(n).reshape(side, side, side)
    rows, cols = [], []
    for d in np.ndindex(3, 3, 3):
        di, dj, dk = (np.asarray(d) - 1).tolist()
        i, j, k = np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")
        ok = (i + di >= 0) & (i + di < side) & (j + dj >= 0) & (j + dj < side) & (k + dk >= 0) & (k + dk < side)
        rows.append(g[i[ok], j[ok], k[ok]])
        cols.append(g[i[ok] + di, j[ok] + dj, k[ok] + dk])
    keys = np.unique(np.concatenate(rows).astype(np.int64) * n + np.concatenate(cols))
    r = keys // n
    indptr = np.concatenate(([0], np.cumsum(np.bincount(r, minlength=n)))).astype(np.int32)
    return indptr, (keys % n).astype(np.int32)


def test_autotuner_times_the_wide_lds_budget_where_the_default_cuts_the_tile():
    """Dense rows (a 27-point stencil): at the default LDS budget the staged tiles of the
    narrow lane groups are cut short, so the autotuner also times the wide budget for them, keeps
    the fastest (lanes, budget) pair and builds the solve's plan with it; the solve then meets
    the tolerance."""

    side, R = 16, 64
    indptr, indices = stencil_27_pattern(side)
    n = side ** 3
    rows = indptr.size - 1
    counts = np.diff(indptr)
    diag_slot = np.array([indptr[i] + np.searchsorted(indices[indptr[i]:indptr[i + 1]], i) for i in range(rows)])
    values = np.full(indices.size, -1.0)
    values[diag_slot] = counts - 1 + 0.05
    solver = adt.solver.BatchedSolver(preconditioner="amg", rows_per_tile=128)
    plan = solver.plan("stencil27", np.array([n]), R, pattern_offsets=(np.zeros(1), np.zeros(1), np.zeros(1)),
                       host_pattern=(indptr, indices))
    (times,) = solver.tuning.values()
    print({f"{k[0]} lanes, {k[1] // 1024} KB": round(v * 1e3, 3) for k, v in times.items()})
    assert (16, native.LDS_BYTES_WIDE) in times and (16, native.LDS_BYTES) in times
    best = min(times, key=times.get)
    assert (plan.lanes, plan.lds_bytes) == best
    rng = np.random.default_rng(4)
    b = torch.zeros((1, n, plan.rp), dtype=torch.float64, device=dev())
    b[0, :, :R] = torch.as_tensor(rng.standard_normal((n, R)), device=dev())
    x = torch.zeros_like(b)
    values_d = torch.as_tensor(values[None], device=dev())
    stats = solver.solve(plan, torch.as_tensor(indptr, device=dev()), torch.as_tensor(indices, device=dev()), values_d,
                         b.reshape(-1), x.reshape(-1))
    assert stats.preconditioner == "amg" and stats.relative_residual.max() <= 1e-10
    solver.close()


def test_iteration_cap_under_the_cycle_runs_exactly_that_many_cycle_iterations():
    """The lane-width autotuner times a fixed number of preconditioned iterations by solving at
    an unreachable tolerance.  A right-hand side that merely reaches max_iter under the cycle
    is reported as not converged; it is not handed to Jacobi PCG (only stalled or broken-down
    ones are), so every right-hand side runs exactly max_iter cycle iterations."""

    rng = np.random.default_rng(5)
    side, R, cap = 20, 13, 12
    indptr, indices, values, _ = random_spd_batch(rng, side)
    S, n = values.shape[0], side * side
    plan = native.PcgPlan(np.full(S, n), R, 16, 32, np.zeros(S), np.zeros(S), np.arange(S) * values.shape[1],
                          host_pattern=(indptr, indices), lds_bytes=8192)
    values_d = torch.as_tensor(values, device=dev())
    plan.attach_amg(values_d, coarse_max=16, stall_limit=64)
    b = torch.as_tensor(rng.standard_normal((S, n, plan.rp)), device=dev())
    b[:, :, R:] = 0.0
    x = torch.zeros_like(b)
    rc, iters, _ = plan.solve(torch.as_tensor(indptr, device=dev()), torch.as_tensor(indices, device=dev()), values_d,
                              b, x, 1e-300, cap, precond=native.PRECOND_AMG, check_every=4)
    stats = plan.stats()
    plan.close()
    assert rc == native.ADT_ERR_NOT_CONVERGED
    assert stats["fallback_rhs"] == 0 and stats["polished_rhs"] == 0
    np.testing.assert_array_equal(iters, np.full_like(iters, cap))


def test_true_residual_restart_brings_every_rhs_to_the_tolerance():
    """At a tight tolerance the CG recurrence residual of some right-hand sides drifts below tol
    while their true residual b - A x is still above it (here 14 of 128, seeded).  Those restart
    once from their iterate; every reported residual is the true one and meets tol, and the
    solutions match a dense solve."""

    rng = np.random.default_rng(13)
    side, S, R, tol = 60, 2, 64, 1e-13
    indptr, indices, values, dense = random_spd_batch(rng, side=side, S=S)
    n = side * side
    plan = native.PcgPlan(np.full(S, n), R, 32, 64, np.zeros(S), np.zeros(S), np.arange(S) * values.shape[1])
    b = rng.standard_normal((S, n, plan.rp))
    b[:, :, R:] = 0.0
    x = torch.zeros((S, n, plan.rp), dtype=torch.float64, device=dev())
    rc, iters, relres = plan.solve(torch.as_tensor(indptr, device=dev()), torch.as_tensor(indices, device=dev()),
                                   torch.as_tensor(values, device=dev()), torch.as_tensor(b, device=dev()), x, tol,
                                   50000)
    restarted = plan.stats()["polished_rhs"]
    plan.close()
    assert rc == 0, native.library().adt_last_error()
    print(f"restarted {restarted} of {S * R}, worst true relative residual {relres.max():.3e}")
    assert restarted > 0
    assert relres.max() <= tol
    xh = x.cpu().numpy()
    for s_ in range(S):
        true = np.linalg.norm(b[s_, :, :R] - dense[s_] @ xh[s_, :, :R], axis=0) / np.linalg.norm(b[s_, :, :R], axis=0)
        # the host's own b - A x at this level carries ~1e-15 |b| of rounding: a few percent of tol
        np.testing.assert_allclose(true, relres[s_], rtol=0.05)
        assert true.max() <= tol * 1.05


@pytest.mark.parametrize("staged", [False, True], ids=["gather", "lds-halo"])
@pytest.mark.parametrize("precond", [native.PRECOND_JACOBI, native.PRECOND_AMG], ids=["jacobi", "amg"])
def test_pcg_solves_random_spd_batches_to_tolerance(staged, precond):
    iters, stats = solve_batch(precond, staged)
    if precond == native.PRECOND_AMG:
        jacobi_iters, _ = solve_batch(native.PRECOND_JACOBI, staged)
        assert stats["levels"] >= 3 and stats["fallback_rhs"] == 0
        assert iters.max() < 0.5 * jacobi_iters.max(), (iters.max(), jacobi_iters.max())


def test_coarsest_operators_beyond_the_lds_size_are_inverted_in_global_memory():
    """A coarsest level of more than 128 rows (900-row grid, coarsening stopped at 400 rows) is
    inverted in global memory instead of LDS; the solutions still match dense solves."""

    rng = np.random.default_rng(2)
    indptr, indices, values, _ = random_spd_batch(rng, 30)
    plan = native.PcgPlan(np.full(2, 900), 13, 16, 32, np.zeros(2), np.zeros(2), np.arange(2) * values.shape[1],
                          host_pattern=(indptr, indices), lds_bytes=8192)
    plan.attach_amg(torch.as_tensor(values, device=dev()), coarse_max=400)
    assert plan.amg_level(plan.amg_levels - 1)["rows"] > 128
    plan.close()
    solve_batch(native.PRECOND_AMG, True, side=30, coarse_max=400)


def test_staged_and_gather_multigrid_iterate_identically():
    """The LDS-staged finest level (with the update fused to the next pre-smoothing and the
    residual double-buffered) performs the same arithmetic in the same order as the gather
    kernels: identical iteration counts for every right-hand side."""

    staged, _ = solve_batch(native.PRECOND_AMG, True, seed=7)
    gather, _ = solve_batch(native.PRECOND_AMG, False, seed=7)
    np.testing.assert_array_equal(staged, gather)


def test_non_finite_preconditioner_hands_over_to_jacobi_pcg():
    """A per-slot damping array of NaN (set on the host) makes every cycle output non-finite:
    flexible CG stops each right-hand side before using it and Jacobi PCG solves all of them
    from their current iterates to the same tolerance."""

    iters, stats = solve_batch(native.PRECOND_LEARNED, True, omega=float("nan"))
    assert stats["fallback_rhs"] == 2 * 13 - 1  # every right-hand side except the zero one


def test_per_slot_damping_equal_to_the_constant_reproduces_the_cycle():
    """The learned entry point with omega_k = the hierarchy's constant damping gives the same
    iterates as the plain multigrid preconditioner (up to the order of one product)."""

    it_amg, _ = solve_batch(native.PRECOND_AMG, True)
    rng = np.random.default_rng(2)
    indptr, indices, values, _ = random_spd_batch(rng)
    plan = native.PcgPlan(np.full(2, 400), 13, 16, 32, np.zeros(2), np.zeros(2), np.arange(2) * values.shape[1],
                          host_pattern=(indptr, indices), lds_bytes=8192)
    plan.attach_amg(torch.as_tensor(values, device=dev()), coarse_max=16)
    omega = plan.amg_level(0)["omega"]
    plan.close()
    it_learned, stats = solve_batch(native.PRECOND_LEARNED, True, omega=omega)
    assert stats["fallback_rhs"] == 0
    assert np.abs(it_learned - it_amg).max() <= 1


@pytest.mark.parametrize("lanes, R", [(8, 5), (64, 70)])
def test_multigrid_flexible_cg_for_other_lane_widths(lanes, R):
    solve_batch(native.PRECOND_AMG, True, side=24, lanes=lanes, R=R, seed=3)


@pytest.mark.parametrize("precond", [native.PRECOND_JACOBI, native.PRECOND_AMG], ids=["jacobi", "amg"])
def test_lane_groups_of_many_tiles_use_segmented_sums(precond):
    """Four-row tiles on 2304 rows give 576 tiles per lane group: the group sums then go through
    the segment pre-reduction; results must still meet the tolerance and match dense solves."""

    solve_batch(precond, True, side=48, lanes=64, R=64, seed=5, rows_per_tile=4)


def test_multigrid_with_tiles_too_tall_for_staging_uses_the_gather_kernels():
    """256-row tiles exceed what the staged kernels hold (8 rows per thread): the finest level
    then runs the gather versions of every cycle kernel; same tolerance and dense solutions."""

    solve_batch(native.PRECOND_AMG, False, side=24, lanes=16, R=13, seed=4, rows_per_tile=256)


def test_multigrid_handles_patterns_of_different_sizes_in_one_batch():
    """Uneven meshes in one plan: two CSR patterns (400 and 169 rows) interleaved over four
    systems, each pattern with its own hierarchy and staging tables; every solution matches a
    dense solve and the multigrid iterates far less than Jacobi PCG."""

    rng = np.random.default_rng(11)
    (pa, ia, va, da), (pb, ib, vb, db) = random_spd_batch(rng, 20), random_spd_batch(rng, 13)
    indptr, indices = np.concatenate([pa, pb]), np.concatenate([ia, ib])
    order = [(0, 0), (1, 0), (0, 1), (1, 1)]  # (pattern, system of that pattern)
    vals = [(va, vb)[q][k] for q, k in order]
    dense = [(da, db)[q][k] for q, k in order]
    rows = np.array([(400, 169)[q] for q, _ in order])
    ptr_off = np.array([(0, pa.size)[q] for q, _ in order])
    idx_off = np.array([(0, ia.size)[q] for q, _ in order])
    val_off = np.concatenate(([0], np.cumsum([v.size for v in vals])[:-1]))
    values_d = torch.as_tensor(np.concatenate(vals), device=dev())
    R = 13
    results = {}
    for precond in (native.PRECOND_JACOBI, native.PRECOND_AMG):
        plan = native.PcgPlan(rows, R, 16, 32, ptr_off, idx_off, val_off, host_pattern=(indptr, indices),
                              lds_bytes=8192)
        assert plan.lds_rows > 0
        if precond == native.PRECOND_AMG:
            assert plan.attach_amg(values_d, coarse_max=16) >= 3
            coarse = [plan.amg_level(1, sys=s)["rows"] for s in range(4)]
            assert coarse[0] == coarse[2] < 400 and coarse[1] == coarse[3] < 169 and coarse[0] != coarse[1]
        b = np.zeros(plan.vec_len)
        sols = []
        for s in range(4):
            bs = rng.standard_normal((rows[s], plan.rp))
            bs[:, R:] = 0.0
            b[plan.vec_off[s]:plan.vec_off[s] + bs.size] = bs.reshape(-1)
            sols.append(np.linalg.solve(dense[s], bs[:, :R]))
        x = torch.zeros(plan.vec_len, dtype=torch.float64, device=dev())
        rc, iters, relres = plan.solve(torch.as_tensor(indptr, device=dev()), torch.as_tensor(indices, device=dev()),
                                       values_d, torch.as_tensor(b, device=dev()), x, 1e-10, 5000, precond=precond)
        assert rc == 0, native.library().adt_last_error()
        assert relres.max() <= 1e-10
        xh = x.cpu().numpy()
        for s in range(4):
            xs = xh[plan.vec_off[s]:plan.vec_off[s] + rows[s] * plan.rp].reshape(rows[s], plan.rp)[:, :R]
            assert relmax(xs, sols[s]) < 1e-7, s
        results[precond] = iters
        plan.close()
    assert results[native.PRECOND_AMG].max() < 0.5 * results[native.PRECOND_JACOBI].max()


@pytest.mark.parametrize("k, Sr, Sc, W", [(4, 64, 64, 1), (4, 96, 96, 1), (4, 64, 64, 3), (10, 17, 40, 3),
                                          (4, 128, 33, 3), (3, 5, 96, 3), (10, 128, 96, 8)],
                         ids=["p1-3d", "p1-96", "p1-3w", "p2-3w", "wide", "narrow", "large"])
@pytest.mark.parametrize("form", ["gram", "default", "block", "block8"])
@pytest.mark.parametrize("C", [47, 48])
def test_cross_gram_sensitivity_matches_the_einsum_reference(k, Sr, Sc, W, form, C, monkeypatch):
    """Jacobian rows from separate receiver and current fields (as in 3D) equal the per-output
    formula with non-symmetric templates, remote indices reading zero and padding columns ignored;
    bitwise repeatable.  Both forms: the per-cell Gram block (matrix cores, four lookups per
    output; forced by ADT_CROSS_FORM=gram) and the cell-block kernel (forced by
    ADT_CROSS_FORM=block: 32 cells and 1024 threads where its LDS fits, else 16/8/4/2 cells;
    ADT_CELL_BLOCK=8 caps it at the narrower widths; a tail block past the last cell).  The
    default takes the cell block at 32 cells or for one field set of linear tetrahedra, and the
    Gram block otherwise; the last case is too large for any cell block and takes the Gram form
    every time.  Odd and even cell counts:
    the 32-cell block stores two cells' outputs as one 16-byte store where the row length is
    even."""

    monkeypatch.setenv("ADT_CROSS_FORM", "block" if form.startswith("block") else form)
    if form == "block8":
        monkeypatch.setenv("ADT_CELL_BLOCK", "8")

    gen = torch.Generator().manual_seed(100 * k + Sr + Sc + W)
    N, D = 80, 39
    phr = torch.randn(W, Sr, N, generator=gen, dtype=torch.float64)
    phc = torch.randn(W, Sc, N, generator=gen, dtype=torch.float64)
    stiff = torch.randn(C, k, k, generator=gen, dtype=torch.float64)
    mass = torch.randn(C, k, k, generator=gen, dtype=torch.float64)
    ws, wm = torch.rand(W, generator=gen, dtype=torch.float64), torch.rand(W, generator=gen, dtype=torch.float64)
    cells = torch.randint(0, N, (C, k), generator=gen)
    a, b = (torch.randint(0, Sc + 1, (D,), generator=gen) for _ in range(2))  # Sc = remote
    m, n = (torch.randint(0, Sr + 1, (D,), generator=gen) for _ in range(2))  # Sr = remote
    pr = torch.cat((phr, torch.zeros_like(phr[:, :1])), dim=1)
    pc = torch.cat((phc, torch.zeros_like(phc[:, :1])), dim=1)
    T = ws[:, None, None, None] * stiff[None] + wm[:, None, None, None] * mass[None]
    expected = -torch.einsum("wdci,wcij,wdcj->dc", (pr[:, m] - pr[:, n])[..., cells], T, (pc[:, a] - pc[:, b])[..., cells])
    rp_r, rp_c = -(-Sr // 16) * 16 + 16, -(-Sc // 8) * 8
    xr = torch.full((W, N, rp_r), 5.0, dtype=torch.float64)  # padding columns must not be read
    xc = torch.full((W, N, rp_c), -3.0, dtype=torch.float64)
    xr[:, :, :Sr] = phr.transpose(1, 2)
    xc[:, :, :Sc] = phc.transpose(1, 2)
    g = lambda t, dt=None: t.to(dev(), dtype=dt or t.dtype).contiguous()  # noqa: E731

    def run():
        out = torch.empty((D, C), dtype=torch.float64, device=dev())
        native.sensitivity_gram_cross(g(xr), np.arange(W) * N * rp_r, rp_r, Sr, g(xc), np.arange(W) * N * rp_c, rp_c,
                                      Sc, *(g(x, torch.int32) for x in (a, b, m, n)), g(cells, torch.int32), g(stiff),
                                      g(mass), ws.numpy(), wm.numpy(), out)
        return out

    first, second = run(), run()
    assert torch.equal(first, second)
    assert relmax(first.cpu(), expected) < 1e-12


@pytest.mark.parametrize("k, S, rp, W", [(3, 9, 16, 7), (4, 40, 64, 7), (10, 128, 128, 7), (3, 96, 96, 7),
                                         (3, 96, 96, 19), (4, 33, 48, 40), (6, 20, 32, 64)])
@pytest.mark.parametrize("with_mass", [True, False])
def test_gram_sensitivity_matches_the_einsum_reference(k, S, rp, W, with_mass):
    """Gram-form Jacobian rows from interleaved (N, Rp) blocks equal the per-output formula
    (sources 9..128, local sizes 3..10, 7..64 wavenumbers: one LDS chunk or several, padding
    columns ignored); bitwise repeatable.  Element templates are positive semidefinite, as every
    stiffness and mass template is, which the Gram form uses; without the mass term the stiffness
    alone is singular (its pivots that vanish to rounding are dropped)."""

    phi, a, b, m, n, cells, _, _, ws, wm = random_problem(k, W=W, S=S, N=90, C=53, D=41, seed=k + S)
    stiff, mass = semidefinite_templates(cells.shape[0], k, seed=k + 3 * S)
    if not with_mass:
        wm = torch.zeros_like(wm)
    problem = (phi, a, b, m, n, cells, stiff, mass, ws, wm)
    expected = einsum_reference(*problem)
    W, _, N = phi.shape
    blocks = torch.zeros((W, N, rp), dtype=torch.float64)
    blocks[:, :, :S] = phi.transpose(1, 2)
    blocks[:, :, S:] = 7.0  # padding columns must not be read
    g = lambda t, dt=None: t.to(dev(), dtype=dt or t.dtype).contiguous()  # noqa: E731

    def run():
        out = torch.empty((a.numel(), cells.shape[0]), dtype=torch.float64, device=dev())
        native.sensitivity_gram(g(blocks), np.arange(W) * N * rp, rp, S, *(g(x, torch.int32) for x in (a, b, m, n)),
                                g(cells, torch.int32), g(stiff), g(mass) if with_mass else None, ws.numpy(),
                                wm.numpy() if with_mass else None, out)
        return out

    out = run()
    assert relmax(out.cpu(), expected) < 1e-12
    assert torch.equal(run(), out)


def column_map(C: int, P: int, seed: int):
    """Seeded cell -> parameter map with dropped cells (-1) and one empty column; the kernels'
    (col_ptr, col_cells) lists and the host reference matrix (C, P)."""

    rng = np.random.default_rng(seed)
    targets = rng.integers(-1, P - 1, C)  # column P - 1 stays empty
    active = np.flatnonzero(targets >= 0)
    order = active[np.argsort(targets[active], kind="stable")]
    ptr = np.concatenate(([0], np.cumsum(np.bincount(targets[active], minlength=P)))).astype(np.int32)
    matrix = np.zeros((C, P))
    matrix[active, targets[active]] = 1.0
    return torch.as_tensor(ptr, device=dev()), torch.as_tensor(order.astype(np.int32), device=dev()), matrix


@pytest.mark.parametrize("k, S", [(3, 19), (10, 40)])
def test_gram_sensitivity_with_fused_column_aggregation(k, S):
    """Aggregated columns (CellAggregation fused into the Gram kernel: each column sums its cells'
    Gram blocks) equal the per-cell einsum reference summed over each column's cells; dropped cells
    and empty columns give nothing; bitwise repeatable."""

    phi, a, b, m, n, cells, _, _, ws, wm = random_problem(k, W=5, S=S, N=80, C=61, D=37, seed=k)
    stiff, mass = semidefinite_templates(cells.shape[0], k, seed=k + 5)
    expected_cells = einsum_reference(phi, a, b, m, n, cells, stiff, mass, ws, wm).numpy()
    col_ptr, col_cells, matrix = column_map(cells.shape[0], 9, seed=k + 1)
    expected = expected_cells @ matrix
    W, _, N = phi.shape
    rp = 16 * ((S + 15) // 16)
    blocks = torch.zeros((W, N, rp), dtype=torch.float64)
    blocks[:, :, :S] = phi.transpose(1, 2)
    g = lambda t, dt=None: t.to(dev(), dtype=dt or t.dtype).contiguous()  # noqa: E731

    def run():
        out = torch.full((a.numel(), 9), float("nan"), dtype=torch.float64, device=dev())
        native.sensitivity_gram(g(blocks), np.arange(W) * N * rp, rp, S, *(g(x, torch.int32) for x in (a, b, m, n)),
                                g(cells, torch.int32), g(stiff), g(mass), ws.numpy(), wm.numpy(), out,
                                columns=(col_ptr, col_cells))
        return out

    out = run()
    assert relmax(out.cpu(), expected) < 1e-12
    assert torch.all(out[:, 8] == 0)
    assert torch.equal(run(), out)


def test_cross_gram_sensitivity_with_fused_column_aggregation():
    rng = np.random.default_rng(9)
    W, Sr, Sc, N, C, D, K = 1, 21, 13, 70, 47, 29, 4
    xr, xc = rng.standard_normal((N, 32)), rng.standard_normal((N, 16))
    cells = rng.integers(0, N, (C, K))
    T = rng.standard_normal((C, K, K))
    a, b = rng.integers(0, Sc + 1, D), rng.integers(0, Sc + 1, D)  # Sc: remote
    m, n = rng.integers(0, Sr + 1, D), rng.integers(0, Sr + 1, D)
    pr = np.vstack((xr[:, :Sr].T, np.zeros((1, N))))
    pc = np.vstack((xc[:, :Sc].T, np.zeros((1, N))))
    rec = (pr[m] - pr[n])[:, cells]
    cur = (pc[a] - pc[b])[:, cells]
    per_cell = -np.einsum("dci,cij,dcj->dc", rec, T, cur)
    col_ptr, col_cells, matrix = column_map(C, 7, seed=3)
    out = torch.empty((D, 7), dtype=torch.float64, device=dev())
    g = lambda x, dt=torch.float64: torch.as_tensor(x, dtype=dt, device=dev()).contiguous()  # noqa: E731
    native.sensitivity_gram_cross(g(xr), [0], 32, Sr, g(xc), [0], 16, Sc, *(g(x, torch.int32) for x in (a, b, m, n)),
                                  g(cells, torch.int32), g(T), None, [1.0], None, out, columns=(col_ptr, col_cells))
    assert relmax(out.cpu().numpy(), per_cell @ matrix) < 1e-12
