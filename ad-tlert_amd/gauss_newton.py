"""Gauss-Newton / Levenberg-Marquardt linear step of the inversion loop on the device.

Follows ``inversion/optimizers.py:128-142, 245-359``: the model increment of the stacked
linearized system ``A dm ~= b`` (weighted log-data Jacobian rows over regularization rows, as
the inversion assembles it) is the least-squares solution by CGLS, with Levenberg-Marquardt
damping appended as ``sqrt(lambda) I`` rows, and is scaled so that no entry exceeds
``max_log_step``.  CGLS runs in the native library (``adt_cgls``): CSR products with a
wavefront per short row or a workgroup per dense Jacobian row, scalars kept on the device.
Only the ``gpu_cgls`` linearized solver is provided (the reference's other choices are SciPy
solvers on the host, which this build does not ship).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import torch

from . import native

LINEARIZED_ALGORITHMS = ("gauss_newton_cgls", "levenberg_marquardt")


@dataclass(frozen=True)
class LinearizedConfig:
    """The settings of the linear step (the matching fields of the reference's InversionConfig,
    ``inversion/config.py:62-80``; any object with these attributes may be passed instead)."""

    optimization_algorithm: str = "gauss_newton_cgls"
    linearized_solver: str = "gpu_cgls"
    lm_damping: float = 1.0e-2
    max_log_step: float | None = 1.0
    cgls_max_iterations: int = 2000
    cgls_tolerance: float = 1.0e-8


def limit_step(delta: np.ndarray, max_step: float | None) -> np.ndarray:
    """``delta`` scaled so that its largest entry is at most ``max_step``."""

    largest = float(np.max(np.abs(delta))) if delta.size else 0.0
    if max_step is None or largest <= max_step:
        return delta
    return delta * (max_step / largest)


def linearized_gradient(matrix, rhs) -> np.ndarray:
    """Gradient of ``||A dm - b||^2 / 2`` at ``dm = 0``: ``-A^T b``."""

    return -np.asarray(matrix.T @ np.asarray(rhs, dtype=float).reshape(-1), dtype=float).reshape(-1)


def _device_csr(matrix: sp.csr_matrix, device):
    return (torch.as_tensor(matrix.indptr.astype(np.int32), device=device),
            torch.as_tensor(matrix.indices.astype(np.int32), device=device),
            torch.as_tensor(matrix.data.astype(np.float64), device=device))


def gpu_cgls(matrix, rhs, *, max_iterations: int, tolerance: float, return_info: bool = False):
    """``argmin ||A x - b||`` by the native CGLS (``optimizers.py:314-359``); A is any scipy
    sparse matrix (or dense array), uploaded with its transpose as CSR."""

    a = sp.csr_matrix(matrix, dtype=np.float64)
    if a.nnz >= 2 ** 31:
        raise ValueError("the linearized system has more than 2^31 - 1 entries")
    device = native.device()
    rhs = np.asarray(rhs, dtype=np.float64).reshape(-1)
    if rhs.shape != (a.shape[0],):
        raise ValueError(f"rhs must have shape ({a.shape[0]},)")
    b = torch.as_tensor(rhs, device=device)
    x = torch.empty(a.shape[1], dtype=torch.float64, device=device)
    iterations, status = native.cgls(_device_csr(a, device), _device_csr(a.T.tocsr(), device), b, x,
                                     int(max_iterations), float(tolerance))
    out = x.cpu().numpy()
    if not np.all(np.isfinite(out)):
        raise ValueError("linearized inversion update contains non-finite values")
    return (out, iterations, native.CGLS_STATUS[status]) if return_info else out


def solve_linearized(matrix, rhs, config) -> np.ndarray:
    """Least-squares solution of ``A x ~= b`` with the configured solver (``optimizers.py:270-311``)."""

    if config.linearized_solver != "gpu_cgls":
        raise NotImplementedError(f"linearized_solver={config.linearized_solver!r}: only the device CGLS "
                                  "('gpu_cgls') is part of this build")
    return gpu_cgls(matrix, rhs, max_iterations=config.cgls_max_iterations, tolerance=config.cgls_tolerance)


def linearized_step(matrix, rhs, current, state, config) -> np.ndarray:
    """Model increment for the stacked linearized system ``A dm ~= b`` (``optimizers.py:245-267``):
    Levenberg-Marquardt appends ``sqrt(lm_damping) I`` rows (and zeros to ``b``); the step is
    limited to ``max_log_step``."""

    del state
    if config.optimization_algorithm not in LINEARIZED_ALGORITHMS:
        raise NotImplementedError(f"optimizer={config.optimization_algorithm!r}: only the linearized algorithms "
                                  f"{LINEARIZED_ALGORITHMS} are part of this build")
    a = sp.csr_matrix(matrix, dtype=np.float64)
    b = np.asarray(rhs, dtype=float).reshape(-1)
    if config.optimization_algorithm == "levenberg_marquardt" and config.lm_damping > 0.0:
        n = int(np.asarray(current).size)
        a = sp.vstack((a, np.sqrt(config.lm_damping) * sp.eye(n, format="csr")), format="csr")
        b = np.concatenate((b, np.zeros(n)))
    return limit_step(solve_linearized(a, b, config), config.max_log_step)
