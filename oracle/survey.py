"""Survey restatement: ABMN slots with a remote electrode and analytic geometric factors.

Follows ``adtlert/survey/core.py:12-120``.  Test infrastructure only.
"""

from __future__ import annotations

import numpy as np

REMOTE = -1


def validate(positions, quads):
    """Validation rules and messages of ``Survey.from_arrays`` (``survey/core.py:34-61``)."""

    positions = np.asarray(positions, dtype=float)
    quads = np.asarray(quads, dtype=np.int64)
    if positions.ndim != 2 or positions.shape[1] not in (2, 3):
        raise ValueError("electrode_positions must have shape (num_electrodes, 2 or 3)")
    if quads.ndim != 2 or quads.shape[1] != 4:
        raise ValueError("measurements must have shape (num_measurements, 4)")
    if np.any(quads < REMOTE):
        raise ValueError("measurements contain negative electrode indices other than -1 (remote)")
    if quads.size and np.any(quads >= positions.shape[0]):
        raise ValueError("measurements reference electrodes outside the survey")
    remote = quads == REMOTE
    if np.any(remote[:, :2].all(axis=1) | remote[:, 2:].all(axis=1)):
        raise ValueError("each current (AB) and potential (MN) pair needs an electrode in the survey")
    distinct = np.where(remote, -1 - np.arange(4), quads)
    if np.any(np.diff(np.sort(distinct, axis=1), axis=1) == 0):
        raise ValueError("each ABMN measurement must use four distinct electrodes")
    return positions, quads


def measurement_slots(quads, electrode_count):
    """Remote electrode mapped to slot ``E`` (``survey/core.py:77-85``)."""

    quads = np.asarray(quads, dtype=np.int64)
    return np.where(quads == REMOTE, electrode_count, quads)


def geometric_factors(positions, quads):
    """``2 pi / (1/AM - 1/AN - 1/BM + 1/BN)``, remote terms dropped (``survey/core.py:87-110``)."""

    positions = np.asarray(positions, dtype=float)
    quads = np.asarray(quads, dtype=np.int64)
    points = positions[np.maximum(quads, 0)]
    real = quads != REMOTE

    def inverse_distance(p, q):
        distance = np.linalg.norm(points[:, p] - points[:, q], axis=-1)
        with np.errstate(divide="ignore"):
            inv = 1.0 / distance
        return np.where(real[:, p] & real[:, q], inv, 0.0)

    return 2.0 * np.pi / (
        inverse_distance(0, 2) - inverse_distance(0, 3) - inverse_distance(1, 2) + inverse_distance(1, 3)
    )


def with_remote_slot(potentials):
    """Zero row and column appended for the remote slot (``survey/core.py:16-20``)."""

    potentials = np.asarray(potentials)
    pad = [(0, 0)] * (potentials.ndim - 2) + [(0, 1), (0, 1)]
    return np.pad(potentials, pad)


def wenner_alpha(electrode_count):
    """Wenner-alpha rows in pyGIMLi 'wa' order (``workflows/terrain.py:303-315``)."""

    if electrode_count < 4:
        raise ValueError("Wenner-alpha surveys need at least four electrodes")
    return np.asarray(
        [
            [start, start + 3 * s, start + s, start + 2 * s]
            for s in range(1, electrode_count // 3 + 1)
            for start in range(electrode_count - 3 * s)
        ],
        dtype=np.int64,
    )
