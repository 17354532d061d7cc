"""IC(0)-preconditioned conjugate gradients on the CPU (FP64): the checker of ``adt_ic0_pcg``.

TEST INFRASTRUCTURE ONLY.  The reference solves the wavenumber systems with a sparse direct
factorization (``forward/cudss.py:63-122``); the device replaces it by preconditioned CG, and this
module restates the IC(0) variant in numpy/scipy so that the device iterates can be compared step
by step:

* ``chow_patel_factor`` -- the fine-grained fixed-point IC(0) factor (Chow and Patel, "Fine-grained
  parallel incomplete LU factorization", SIAM J. Sci. Comput. 37 (2015) C169-C193): every entry of
  L on the lower pattern of A is recomputed from the previous sweep's values,
  ``l_ij = (a_ij - sum_{k<j} l_ik l_jk) / l_jj``, ``l_ii = sqrt(a_ii - sum_{k<i} l_ik^2)``, from
  ``l_ij = a_ij / sqrt(a_jj)``.  Pinned against ``exact_ic0`` (the sequential textbook IC(0), Golub
  and Van Loan 4th ed. section 11.5.8): the fixed point is that factor, and with at least n sweeps
  the iteration reaches it exactly (its dependency graph has depth < n);
* ``apply`` -- ``z = M^-1 r`` with ``1 + sweeps`` Jacobi sweeps of ``L y = r`` and of ``L^T z = y``.
  Pinned against exact triangular solves (``scipy.linalg.solve_triangular``) for ``sweeps >= n - 1``
  and for symmetry of the operator at any sweep count;
* ``pcg`` -- preconditioned CG (Hestenes and Stiefel 1952) per column, ``||r|| <= tol ||b||``, with the
  same restart on the true residual as the device.  Pinned against ``numpy.linalg.solve``.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def lower_pattern(matrix: sp.csr_matrix) -> sp.csr_matrix:
    """The lower triangle of ``matrix`` (diagonal included) as sorted CSR."""

    low = sp.tril(sp.csr_matrix(matrix), format="csr")
    low.sort_indices()
    return low


def exact_ic0(matrix) -> sp.csr_matrix:
    """Sequential IC(0): the Cholesky recurrences restricted to the lower pattern of ``matrix``."""

    low = lower_pattern(matrix)
    n = low.shape[0]
    rows = [dict(zip(low.indices[low.indptr[i]:low.indptr[i + 1]].tolist(),
                     low.data[low.indptr[i]:low.indptr[i + 1]].tolist())) for i in range(n)]
    out = [dict() for _ in range(n)]
    for i in range(n):
        for j in sorted(rows[i]):
            s = sum(out[i][k] * out[j][k] for k in out[i] if k < j and k in out[j])
            if i == j:
                out[i][i] = np.sqrt(rows[i][i] - s)
            else:
                out[i][j] = (rows[i][j] - s) / out[j][j]
    data = np.array([out[i][j] for i in range(n) for j in sorted(rows[i])])
    return sp.csr_matrix((data, low.indices.copy(), low.indptr.copy()), shape=low.shape)


def chow_patel_factor(matrix, sweeps: int) -> sp.csr_matrix:
    """``sweeps`` Jacobi-style fixed-point sweeps of the IC(0) entries (all from the previous sweep)."""

    low = lower_pattern(matrix)
    n = low.shape[0]
    ptr, idx, a = low.indptr, low.indices, low.data
    diag = a[ptr[1:] - 1]  # sorted rows: the diagonal is each row's last entry
    rows = np.repeat(np.arange(n), np.diff(ptr))
    cur = np.where(rows == idx, np.sqrt(np.abs(a)), a / np.sqrt(np.abs(diag[idx])))
    for _ in range(sweeps):
        new = np.empty_like(cur)
        for e in range(idx.size):
            i, j = rows[e], idx[e]
            ri = dict(zip(idx[ptr[i]:ptr[i + 1] - 1].tolist(), cur[ptr[i]:ptr[i + 1] - 1].tolist()))
            rj = dict(zip(idx[ptr[j]:ptr[j + 1] - 1].tolist(), cur[ptr[j]:ptr[j + 1] - 1].tolist()))
            s = sum(v * rj[k] for k, v in ri.items() if k < j and k in rj)
            if i == j:
                d = a[e] - s
                new[e] = np.sqrt(d) if d > 0 else np.sqrt(abs(a[e]))
            else:
                new[e] = (a[e] - s) / cur[ptr[j + 1] - 1]
        cur = new
    return sp.csr_matrix((cur, idx.copy(), ptr.copy()), shape=low.shape)


def apply(factor: sp.csr_matrix, r: np.ndarray, sweeps: int) -> np.ndarray:
    """``z = P^T P r``: ``P r`` by ``1 + sweeps`` Jacobi sweeps of ``L y = r`` from ``y = D^-1 r``,
    then ``P^T y`` by as many sweeps of ``L^T z = y``."""

    d = factor.diagonal()
    strict = sp.tril(factor, k=-1, format="csr")
    y = r / d[:, None] if r.ndim == 2 else r / d
    dd = d[:, None] if r.ndim == 2 else d
    for _ in range(sweeps):
        y = (r - strict @ y) / dd
    z = y / dd
    upper = strict.T.tocsr()
    for _ in range(sweeps):
        z = (y - upper @ z) / dd
    return z


def pcg(matrix, rhs: np.ndarray, x0: np.ndarray | None = None, *, tol: float = 1e-10, max_iter: int = 10000,
        factor_sweeps: int = 3, solve_sweeps: int = 2):
    """IC(0)-PCG per column of ``rhs`` ``(n, r)``; returns ``(x, iterations (r,))``.  Each column
    stops at ``||r|| <= tol ||b||`` (recurrence residual), then is checked on the true residual and
    continued from its iterate (at most two restarts); ``max_iter`` bounds the total."""

    a = sp.csr_matrix(matrix)
    factor = chow_patel_factor(a, factor_sweeps)
    b = np.asarray(rhs, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    iters = np.zeros(b.shape[1], dtype=np.int64)
    for c in range(b.shape[1]):
        bb = float(b[:, c] @ b[:, c])
        if bb == 0.0:
            continue
        done = 0
        for _restart in range(3):
            r = b[:, c] - a @ x[:, c]
            if float(r @ r) <= tol * tol * bb:
                break
            z = apply(factor, r, solve_sweeps)
            p = z.copy()
            rz = float(r @ z)
            while done < max_iter:
                q = a @ p
                alpha = rz / float(p @ q)
                x[:, c] += alpha * p
                r -= alpha * q
                done += 1
                if float(r @ r) <= tol * tol * bb:
                    break
                z = apply(factor, r, solve_sweeps)
                rz_new = float(r @ z)
                p = z + (rz_new / rz) * p
                rz = rz_new
            if done >= max_iter:
                break
        iters[c] = done
    return x, iters
