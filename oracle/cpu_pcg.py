"""Block Jacobi-preconditioned conjugate gradients on the CPU (FP64), many right-hand sides.

TEST INFRASTRUCTURE ONLY (the timed CPU baseline of the 3D workload, where the oracle's sparse
direct factorization takes minutes).  It is the textbook method (Hestenes and Stiefel 1952)
with the diagonal preconditioner, run on all right-hand sides at once: one sparse-times-dense
product per iteration (torch's CPU CSR kernel, which uses every intra-op thread), column-wise
inner products, and each right-hand side frozen once ``||b - A x|| <= tol ||b||`` by its
recurrence residual.  It is the CPU counterpart of the GPU's plain Jacobi-PCG and is pinned
against ``scipy.sparse.linalg.splu`` in ``tests/test_oracle.py``.
"""

from __future__ import annotations

import warnings

import numpy as np
import torch


def jacobi_pcg(matrix, rhs: np.ndarray, tol: float = 1e-10, max_iter: int = 20000, threads: int | None = None):
    """Solve ``A X = B`` for the columns of ``rhs`` ``(N, R)``; returns ``(X, iterations (R,))``.

    ``matrix`` is a scipy CSR matrix (symmetric positive definite); ``threads`` sets torch's
    intra-op thread count for the call."""

    previous = torch.get_num_threads()
    if threads:
        torch.set_num_threads(int(threads))
    try:
        with warnings.catch_warnings():
            warnings.filterwarnings("ignore", message="Sparse CSR tensor support is in beta")
            a = torch.sparse_csr_tensor(torch.as_tensor(matrix.indptr, dtype=torch.int64),
                                        torch.as_tensor(matrix.indices, dtype=torch.int64),
                                        torch.as_tensor(matrix.data, dtype=torch.float64), size=matrix.shape)
        b = torch.as_tensor(np.ascontiguousarray(rhs), dtype=torch.float64)
        dinv = 1.0 / torch.as_tensor(matrix.diagonal(), dtype=torch.float64)[:, None]
        x = torch.zeros_like(b)
        r = b.clone()
        z = dinv * r
        p = z.clone()
        rz = (r * z).sum(0)
        limit = (tol * tol) * (b * b).sum(0)
        iterations = torch.zeros(b.shape[1], dtype=torch.int64)
        active = (r * r).sum(0) > limit
        for _ in range(max_iter):
            if not bool(active.any()):
                break
            q = a @ p
            alpha = torch.where(active, rz / (p * q).sum(0), torch.zeros_like(rz))
            x += alpha * p
            r -= alpha * q
            iterations += active.to(torch.int64)
            active &= (r * r).sum(0) > limit
            z = dinv * r
            rz_next = (r * z).sum(0)
            beta = torch.where(active, rz_next / rz, torch.zeros_like(rz))
            p = z + beta * p
            rz = rz_next
        return x.numpy(), iterations.numpy()
    finally:
        torch.set_num_threads(previous)
