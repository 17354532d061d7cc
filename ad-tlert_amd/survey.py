"""ABMN survey: electrode positions, measurement slots with a remote electrode, geometric factors.

Same interface and error behaviour as ``adtlert/survey/core.py:12-120``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from .dtypes import FLOAT_DTYPE, INT_DTYPE

REMOTE_ELECTRODE = -1
"""ABMN index of an electrode at infinity (pole arrays); its potential is zero."""


def with_remote_slot(potentials: torch.Tensor) -> torch.Tensor:
    """Append the zero row and column of the remote slot to ``(..., sources, receivers)``."""

    return torch.nn.functional.pad(potentials, (0, 1, 0, 1))


@dataclass(frozen=True)
class Survey:
    """Electrode geometry and ABMN measurement indexing (``survey/core.py:23-120``)."""

    electrode_positions: torch.Tensor
    measurements: torch.Tensor

    @classmethod
    def from_arrays(cls, electrode_positions, measurements) -> "Survey":
        positions = torch.as_tensor(electrode_positions, dtype=FLOAT_DTYPE)
        quads = torch.as_tensor(measurements, dtype=INT_DTYPE)
        if positions.ndim != 2 or positions.shape[1] not in (2, 3):
            raise ValueError("electrode_positions must have shape (num_electrodes, 2 or 3)")
        if quads.ndim != 2 or quads.shape[1] != 4:
            raise ValueError("measurements must have shape (num_measurements, 4)")
        if bool(torch.any(quads < REMOTE_ELECTRODE)):
            raise ValueError(f"measurements contain negative electrode indices other than {REMOTE_ELECTRODE} (remote)")
        if quads.numel() and bool(torch.any(quads >= positions.shape[0])):
            raise ValueError("measurements reference electrodes outside the survey")
        remote = quads == REMOTE_ELECTRODE
        if bool(torch.any(remote[:, :2].all(dim=1) | remote[:, 2:].all(dim=1))):
            raise ValueError("each current (AB) and potential (MN) pair needs an electrode in the survey")
        distinct = torch.where(remote, -1 - torch.arange(4, dtype=quads.dtype), quads)
        if bool(torch.any(torch.diff(torch.sort(distinct, dim=1).values, dim=1) == 0)):
            raise ValueError("each ABMN measurement must use four distinct electrodes")
        return cls(electrode_positions=positions, measurements=quads)

    @property
    def electrode_count(self) -> int:
        return int(self.electrode_positions.shape[0])

    @property
    def measurement_count(self) -> int:
        return int(self.measurements.shape[0])

    @property
    def dimension(self) -> int:
        return int(self.electrode_positions.shape[1])

    def measurement_slots(self) -> torch.Tensor:
        """ABMN with the remote electrode mapped to slot ``electrode_count``."""

        quads = self.measurements.long()
        return torch.where(quads == REMOTE_ELECTRODE, self.electrode_count, quads)

    def geometric_factors(self) -> torch.Tensor:
        """``2 pi / (1/AM - 1/AN - 1/BM + 1/BN)``; terms with a remote electrode vanish."""

        quads = self.measurements.long()
        points = self.electrode_positions[quads.clamp_min(0)]
        real = quads != REMOTE_ELECTRODE

        def inverse_distance(p: int, q: int) -> torch.Tensor:
            distance = torch.linalg.norm(points[:, p] - points[:, q], dim=-1)
            return torch.where(real[:, p] & real[:, q], 1.0 / distance, 0.0)

        return 2.0 * math.pi / (inverse_distance(0, 2) - inverse_distance(0, 3) - inverse_distance(1, 2)
                                + inverse_distance(1, 3))

    def apparent_resistivity(self, voltages, currents=1.0) -> torch.Tensor:
        voltages = torch.as_tensor(voltages, dtype=FLOAT_DTYPE)
        return self.geometric_factors() * voltages / torch.as_tensor(currents, dtype=FLOAT_DTYPE)
