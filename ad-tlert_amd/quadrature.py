"""Inverse cosine-transform quadrature for the 2.5D wavenumber integral.

Same rule as ``adtlert/forward/integration.py:23-49``: Gauss-Legendre in ``sqrt(k)`` on
``[0, k0]`` with ``k0 = 1/(2 r_min)`` and ``max(int(6 log10(r_max/r_min)), 4)`` points, plus a
4-point Gauss-Laguerre tail, so that ``sum_w weights_w f(k_w) ~ (1/pi) int_0^inf f(k) dk``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch
from scipy.special import roots_laguerre, roots_legendre

from .dtypes import FLOAT_DTYPE


@dataclass(frozen=True)
class CosineTransformWeights:
    wavenumbers: torch.Tensor
    weights: torch.Tensor

    def __iter__(self):
        return iter((self.wavenumbers, self.weights))


def survey_wavenumber_bounds(survey) -> tuple[float, float]:
    """Half the smallest and twice the largest electrode separation."""

    positions = np.asarray(survey.electrode_positions, dtype=float)
    distances = np.linalg.norm(positions[:, None, :] - positions[None, :, :], axis=-1)
    positive = distances[distances > 0.0]
    return float(np.min(positive) / 2.0), float(np.max(positive) * 2.0)


def build_inverse_cosine_weights(r_min: float, r_max: float) -> CosineTransformWeights:
    k0 = 1.0 / (2.0 * r_min)
    x, w = roots_legendre(max(int(6.0 * np.log10(r_max / r_min)), 4))
    x, w = 0.5 * (x + 1.0), 0.5 * w
    tx, tw = roots_laguerre(4)
    wavenumbers = np.concatenate((k0 * x * x, k0 * (tx + 1.0)))
    weights = np.concatenate((2.0 * k0 * x * w / np.pi, k0 * np.exp(tx) * tw / np.pi))
    return CosineTransformWeights(torch.as_tensor(wavenumbers, dtype=FLOAT_DTYPE),
                                  torch.as_tensor(weights, dtype=FLOAT_DTYPE))
