"""Inverse cosine-transform quadrature of the 2.5D wavenumber integral.

Follows ``adtlert/forward/integration.py:23-49``.  Test infrastructure only.
"""

from __future__ import annotations

import numpy as np
from scipy.special import roots_laguerre, roots_legendre


def wavenumber_bounds(positions):
    """Half the smallest and twice the largest electrode separation (``integration.py:23-29``)."""

    positions = np.asarray(positions, dtype=float)
    distances = np.linalg.norm(positions[:, None, :] - positions[None, :, :], axis=-1)
    positive = distances[distances > 0.0]
    return float(np.min(positive) / 2.0), float(np.max(positive) * 2.0)


def inverse_cosine_weights(r_min, r_max):
    """Gauss-Legendre in sqrt(k) on [0, k0] plus a 4-point Gauss-Laguerre tail (``integration.py:32-49``).

    Returns ``(wavenumbers, weights)`` in float64 with ``sum_w weights_w f(k_w) ~ (1/pi) int_0^inf f``.
    """

    k0 = 1.0 / (2.0 * r_min)
    count = max(int(6.0 * np.log10(r_max / r_min)), 4)
    x, w = roots_legendre(count)
    x, w = 0.5 * (x + 1.0), 0.5 * w
    tx, tw = roots_laguerre(4)
    wavenumbers = np.concatenate((k0 * x * x, k0 * (tx + 1.0)))
    weights = np.concatenate((2.0 * k0 * x * w / np.pi, k0 * np.exp(tx) * tw / np.pi))
    return wavenumbers, weights
