"""ABMN surveys: electrode positions, measurement slots with a remote electrode, geometric factors.

Interface of ``adtlert/survey/core.py:12-120``: ``Survey.from_arrays`` raises ``ValueError``
with the reference's messages (callers and tests match on them), ``measurement_slots`` maps
the remote electrode to one slot past the last electrode, and ``geometric_factors`` is the
analytic half-space factor ``2 pi / (1/AM - 1/AN - 1/BM + 1/BN)`` with remote terms dropped.

Geometric factors are formed here from per-slot positions in which the remote slot holds NaN:
every inverse distance involving it is NaN and is replaced by zero, so the four terms need no
case analysis.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from .dtypes import FLOAT_DTYPE, INT_DTYPE

REMOTE_ELECTRODE = -1
"""ABMN index of an electrode at infinity (pole arrays); its potential is zero."""


def with_remote_slot(potentials: torch.Tensor) -> torch.Tensor:
    """``(..., sources, receivers)`` potentials with a zero last row and column: the remote slot."""

    return torch.nn.functional.pad(potentials, (0, 1, 0, 1))


def _validation_error(positions: torch.Tensor, quads: torch.Tensor) -> str | None:
    """First rule of ``Survey.from_arrays`` that the arrays break, or None."""

    if positions.dim() != 2 or int(positions.shape[-1]) not in (2, 3):
        return "electrode_positions must have shape (num_electrodes, 2 or 3)"
    if quads.dim() != 2 or int(quads.shape[-1]) != 4:
        return "measurements must have shape (num_measurements, 4)"
    if quads.numel() == 0:
        return None
    if int(quads.min()) < REMOTE_ELECTRODE:
        return f"measurements contain negative electrode indices other than {REMOTE_ELECTRODE} (remote)"
    if int(quads.max()) >= int(positions.shape[0]):
        return "measurements reference electrodes outside the survey"
    remote = quads == REMOTE_ELECTRODE
    current_at_infinity = remote[:, 0] & remote[:, 1]
    potential_at_infinity = remote[:, 2] & remote[:, 3]
    if bool((current_at_infinity | potential_at_infinity).any()):
        return "each current (AB) and potential (MN) pair needs an electrode in the survey"
    # remote entries become distinct negative labels, so only real electrodes can collide
    labels = torch.where(remote, -2 - torch.arange(4, dtype=quads.dtype), quads)
    collide = torch.zeros(quads.shape[0], dtype=torch.bool)
    for i in range(4):
        for j in range(i + 1, 4):
            collide |= labels[:, i] == labels[:, j]
    if bool(collide.any()):
        return "each ABMN measurement must use four distinct electrodes"
    return None


@dataclass(frozen=True)
class Survey:
    """Electrode coordinates ``(E, 2 or 3)`` and ABMN rows ``(D, 4)``; ``REMOTE_ELECTRODE``
    places B and/or N at infinity (pole-dipole, dipole-pole, pole-pole)."""

    electrode_positions: torch.Tensor
    measurements: torch.Tensor

    @classmethod
    def from_arrays(cls, electrode_positions, measurements) -> "Survey":
        positions = torch.as_tensor(electrode_positions, dtype=FLOAT_DTYPE)
        quads = torch.as_tensor(measurements, dtype=INT_DTYPE)
        problem = _validation_error(positions, quads)
        if problem is not None:
            raise ValueError(problem)
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
        """``(D, 4)`` int64 slots: electrode indices, the remote electrode as ``electrode_count``."""

        slots = self.measurements.to(torch.int64)
        return slots.masked_fill(slots == REMOTE_ELECTRODE, self.electrode_count)

    def geometric_factors(self) -> torch.Tensor:
        """Analytic half-space factor ``2 pi / (1/AM - 1/AN - 1/BM + 1/BN)``, remote terms zero."""

        remote_row = torch.full((1, self.dimension), float("nan"), dtype=self.electrode_positions.dtype)
        table = torch.cat((self.electrode_positions, remote_row))  # slot E: the electrode at infinity
        at = table[self.measurement_slots()]  # (D, 4, dim): A, B, M, N
        inverse = 1.0 / torch.linalg.norm(at[:, :2, None, :] - at[:, None, 2:, :], dim=-1)  # (D, [A, B], [M, N])
        inverse = torch.where(torch.isnan(inverse), torch.zeros_like(inverse), inverse)
        coupling = (inverse[:, 0, 0] - inverse[:, 0, 1]) - (inverse[:, 1, 0] - inverse[:, 1, 1])
        return 2.0 * math.pi / coupling

    def apparent_resistivity(self, voltages, currents=1.0) -> torch.Tensor:
        """``k * V / I``."""

        k = self.geometric_factors()
        return k * torch.as_tensor(voltages, dtype=FLOAT_DTYPE) / torch.as_tensor(currents, dtype=FLOAT_DTYPE)
