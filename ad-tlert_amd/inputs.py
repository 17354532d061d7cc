"""Argument checks shared by the 2.5D and 3D operators.

The exception types and message texts are the reference's contract
(``adtlert/forward/inputs.py:13-100``): callers and the reference's tests match on them.  Every
check converts its input once to a tensor (or an int64 array for parameter maps) and then
tests it as a whole.
"""

from __future__ import annotations

import numpy as np
import torch

from .dtypes import FLOAT_DTYPE

Tensor = torch.Tensor

_NOT_POSITIVE = "conductivity must contain positive finite values"


def _finite(t: Tensor) -> bool:
    return bool(torch.isfinite(t).all())


def _per_entry(values, count: int, dtype: torch.dtype, shape_message: str) -> Tensor:
    """``values`` as a ``(count,)`` tensor: a 0-d input is broadcast (a view), anything but
    ``(count,)`` is rejected with ``shape_message``."""

    t = torch.as_tensor(values, dtype=dtype)
    if t.dim() == 0:
        return t.expand(count)
    if tuple(t.shape) != (count,):
        raise ValueError(shape_message)
    return t


def check_finite(values: Tensor, context: str) -> None:
    """``FloatingPointError`` when a computed result holds NaN or infinity."""

    if not _finite(values):
        raise FloatingPointError(f"{context} produced non-finite values")


def cell_vector(values, count: int, dtype: torch.dtype = FLOAT_DTYPE) -> Tensor:
    """Finite per-cell values (conductivity directions); scalars broadcast."""

    vector = _per_entry(values, count, dtype, f"conductivity must be scalar or shape ({count},)")
    if not _finite(vector):
        raise ValueError("conductivity must contain finite values")
    return vector


def conductivity_model(values, count: int, dtype: torch.dtype = FLOAT_DTYPE) -> Tensor:
    """One conductivity model: finite and strictly positive per cell."""

    model = cell_vector(values, count, dtype)
    if not bool((model > 0).all()):
        raise ValueError(_NOT_POSITIVE)
    return model


def conductivity_series(values, count: int, dtype: torch.dtype = FLOAT_DTYPE) -> Tensor:
    """A series of models ``(n_steps, count)``, each finite and strictly positive."""

    models = torch.as_tensor(values, dtype=dtype)
    if models.dim() != 2 or int(models.shape[1]) != count:
        raise ValueError(f"conductivities must have shape (n_steps, {count})")
    if not bool((torch.isfinite(models) & (models > 0)).all()):
        raise ValueError(_NOT_POSITIVE)
    return models


def measurement_vector(values, count: int, dtype: torch.dtype) -> Tensor:
    """A per-measurement cotangent; scalars broadcast."""

    return _per_entry(values, count, dtype, f"cotangent must be scalar or shape ({count},)")


def cotangent_series(values, steps: int, count: int) -> Tensor:
    """Per-step cotangents of a model series, exactly ``(steps, count)``."""

    series = torch.as_tensor(values, dtype=FLOAT_DTYPE)
    expected = (steps, count)
    if tuple(series.shape) != expected:
        raise ValueError(f"cotangents must have shape {expected}")
    return series


def parameter_targets(cell_parameter_ids, parameter_count, cell_count: int) -> tuple[np.ndarray, int]:
    """Output column of every cell (``-1``: the cell is dropped) and the number of columns.

    Without a map every cell is its own column.  A map needs at least one active cell, ids of
    ``-1`` or more, and a ``parameter_count`` (if given) above its largest id."""

    if cell_parameter_ids is None:
        if parameter_count is not None:
            raise ValueError("parameter_count requires cell_parameter_ids")
        return np.arange(cell_count, dtype=np.int64), cell_count
    raw = cell_parameter_ids.detach().cpu().numpy() if isinstance(cell_parameter_ids, Tensor) else cell_parameter_ids
    ids = np.asarray(raw).astype(np.int64).ravel()
    if ids.size != cell_count:
        raise ValueError(f"cell_parameter_ids must have shape ({cell_count},)")
    largest = int(ids.max(initial=-1))
    if int(ids.min(initial=0)) < -1 or largest < 0:
        raise ValueError("cell_parameter_ids may only contain -1 or non-negative ids, with at least one active")
    count = largest + 1 if parameter_count is None else int(parameter_count)
    if count <= largest:
        raise ValueError("parameter_count is smaller than the largest cell parameter id")
    return ids, count


def option(value, choices: tuple[str, ...], name: str) -> str:
    """Case- and whitespace-insensitive choice among ``choices``."""

    key = str(value).strip().lower()
    if key in choices:
        return key
    raise ValueError(f"{name} must be one of: {', '.join(sorted(choices))}")
