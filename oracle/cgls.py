"""CGLS for ``min ||A x - b||_2`` in FP64 numpy (TEST INFRASTRUCTURE ONLY).

Restates the reference's GPU CGLS (``inversion/optimizers.py:314-359``): the same iteration
(Hestenes-Stiefel on the normal equations without forming ``A^T A``) and the same stop rule,
``gamma' <= 0`` or ``gamma' / gamma0 < tol`` with ``gamma = ||A^T r||^2``; also stops when
``||A p||^2 <= 0``.  Returns the solution and the number of completed iterations.  Pinned in
``tests/test_oracle.py`` against ``numpy.linalg.lstsq`` and the closed-form damped normal
equations.
"""

from __future__ import annotations

import numpy as np


def cgls(matrix, rhs, max_iterations: int, tolerance: float):
    a = matrix
    r = np.asarray(rhs, dtype=float).reshape(-1).copy()
    x = np.zeros(a.shape[1])
    s = np.asarray(a.T @ r, dtype=float).reshape(-1)
    p = s.copy()
    gamma = float(s @ s)
    gamma0 = gamma
    if gamma0 <= 0.0 or float(r @ r) <= 0.0:
        return x, 0
    done = 0
    for _ in range(int(max_iterations)):
        q = np.asarray(a @ p, dtype=float).reshape(-1)
        den = float(q @ q)
        if den <= 0.0:
            break
        alpha = gamma / den
        x = x + alpha * p
        r = r - alpha * q
        done += 1
        s = np.asarray(a.T @ r, dtype=float).reshape(-1)
        gamma_new = float(s @ s)
        if gamma_new <= 0.0 or gamma_new / gamma0 < tolerance:
            break
        p = s + (gamma_new / gamma) * p
        gamma = gamma_new
    return x, done
