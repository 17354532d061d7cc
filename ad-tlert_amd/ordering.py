"""Locality-preserving DOF numbering for the device-side operators.

The multi-RHS sparse products gather every neighbour row of a node several times.  In the
natural row-by-row numbering of a graded tensor mesh those re-reads are about two grid rows
apart, which overflows L2 once hundreds of workgroups run.  Numbering the DOFs along a
Hilbert curve makes each contiguous row range (one workgroup tile) a compact patch whose
neighbours are mostly inside the patch, so the gathers hit L2.  The permutation is internal:
fields handed back to callers are returned in the mesh's own node order.
"""

from __future__ import annotations

import os

import numpy as np


def _hilbert_index(ints: np.ndarray, bits: int) -> np.ndarray:
    """Hilbert-curve index of integer points ``(n, d)`` with ``bits`` bits per axis
    (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004), vectorized."""

    x = ints.astype(np.uint64).copy()
    n, d = x.shape
    m = np.uint64(1) << np.uint64(bits - 1)
    # inverse undo excess work
    q = m
    while q > 1:
        p = q - np.uint64(1)
        for i in range(d):
            hit = (x[:, i] & q) != 0
            x[hit, 0] ^= p
            t = (x[~hit, 0] ^ x[~hit, i]) & p
            x[~hit, 0] ^= t
            x[~hit, i] ^= t
        q >>= np.uint64(1)
    # Gray encode
    for i in range(1, d):
        x[:, i] ^= x[:, i - 1]
    t = np.zeros(n, dtype=np.uint64)
    q = m
    while q > 1:
        hit = (x[:, d - 1] & q) != 0
        t[hit] ^= q - np.uint64(1)
        q >>= np.uint64(1)
    for i in range(d):
        x[:, i] ^= t
    # interleave the transposed bits into one index (most significant first)
    index = np.zeros(n, dtype=np.uint64)
    for b in range(bits - 1, -1, -1):
        for i in range(d):
            index = (index << np.uint64(1)) | ((x[:, i] >> np.uint64(b)) & np.uint64(1))
    return index


def hilbert_order(points: np.ndarray, bits: int | None = None) -> np.ndarray:
    """Permutation ``perm`` (new -> old) sorting points along a Hilbert curve of their bounding box."""

    points = np.asarray(points, dtype=float)
    n, d = points.shape
    if n <= 2:
        return np.arange(n)
    if bits is None:
        bits = 21 if d == 3 else 31
    lo, hi = points.min(axis=0), points.max(axis=0)
    span = np.where(hi - lo > 0, hi - lo, 1.0)
    scale = float((1 << bits) - 1)
    ints = np.floor((points - lo) / span * scale).astype(np.int64)
    return np.argsort(_hilbert_index(ints, bits), kind="stable")


def inverse(perm: np.ndarray) -> np.ndarray:
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    return inv


def rank_coordinates(points: np.ndarray) -> np.ndarray:
    """Each coordinate replaced by its rank among the distinct values on its axis.

    On a graded tensor mesh this is the node's grid index, so a Hilbert curve over it cuts tiles
    that are compact in the mesh graph rather than in physical space (where a 0.5 m vertical and
    1.25 m lateral spacing makes physically cubic patches span 2.5 times more layers than
    columns).  On unstructured meshes it equalizes the node density per axis."""

    points = np.asarray(points, dtype=float)
    out = np.empty_like(points)
    for d in range(points.shape[1]):
        _, rank = np.unique(points[:, d], return_inverse=True)
        out[:, d] = rank
    return out


def dof_order(points: np.ndarray) -> np.ndarray:
    """Device DOF numbering: Hilbert order of the rank coordinates (default), of the physical
    coordinates (``ADT_DOF_ORDER=hilbert-physical``) or the mesh's own (``ADT_DOF_ORDER=natural``)."""

    mode = os.environ.get("ADT_DOF_ORDER", "hilbert").strip().lower()
    if mode == "natural":
        return np.arange(np.asarray(points).shape[0])
    if mode == "hilbert-physical":
        return hilbert_order(points)
    return hilbert_order(rank_coordinates(points))
