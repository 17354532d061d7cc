"""PyTorch autograd bridge for the 2.5D and 3D operators.

Contract of ``adtlert/forward/autograd.py:17-207``:

* ``apparent_resistivity_autograd(conductivity, forward_operator, currents=1.0)`` maps one model
  ``(n_cells,)`` or a series ``(n_steps, n_cells)`` to apparent resistivities ``(n_data,)`` or
  ``(n_steps, n_data)`` on the conductivity's device and dtype;
* reverse mode goes through the operator's ``vjp_series`` (the native adjoint kernels), forward
  mode through ``jvp`` per step; conductivity and currents (a scalar, a length-one tensor or one
  value per datum) both receive derivatives; the solve is opaque, so there are no second
  derivatives;
* inputs are validated with the reference's exception types and messages.

The work runs on the host copies in ``FLOAT_DTYPE`` that the operators take.  With
``rho_a = k R(sigma) / I``: ``d rho_a = k dR / I - rho_a dI / I``.
"""

from __future__ import annotations

from typing import Any

import torch
from torch.autograd.function import once_differentiable

from .dtypes import FLOAT_DTYPE
from .ert2p5d import ERTForward2p5D
from .ert3d import ERTForward3D

_OPERATORS = (ERTForward2p5D, ERTForward3D)


def _host(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to(device="cpu", dtype=FLOAT_DTYPE)


def _check(conductivity, currents, forward_operator) -> None:
    """Raise the reference's error for the first invalid argument."""

    if not isinstance(forward_operator, _OPERATORS):
        raise TypeError("forward_operator must be an ERTForward2p5D or ERTForward3D instance")
    floating = lambda t: isinstance(t, torch.Tensor) and t.is_floating_point()  # noqa: E731
    if not floating(conductivity):
        raise TypeError("conductivity must be a floating-point torch.Tensor")
    cells = forward_operator.mesh.cell_count
    if conductivity.dim() not in (1, 2) or int(conductivity.shape[-1]) != cells:
        raise ValueError(f"conductivity must have shape ({cells},) or (n_steps, {cells}), "
                         f"got {tuple(conductivity.shape)}")
    if not bool((torch.isfinite(conductivity) & (conductivity > 0)).all()):
        raise ValueError("conductivity must contain finite positive values")
    if not floating(currents):
        raise TypeError("currents must be a floating-point torch.Tensor")
    data = forward_operator.survey.measurement_count
    if tuple(currents.shape) not in ((), (1,), (data,)):
        raise ValueError(f"currents must be scalar or have shape ({data},), got {tuple(currents.shape)}")
    if not bool((torch.isfinite(currents) & (currents != 0)).all()):
        raise ValueError("currents must contain finite non-zero values")


def _steps(t: torch.Tensor, width: int) -> torch.Tensor:
    """A model or a series as ``(n_steps, width)`` rows."""

    return t.reshape(-1, width)


class ERT2p5DApparentResistivityFunction(torch.autograd.Function):
    """Differentiable apparent resistivity of a 2.5D or 3D operator (one model or a series)."""

    @staticmethod
    def forward(ctx: Any, conductivity, currents, forward_operator):
        _check(conductivity, currents, forward_operator)
        cells = conductivity.shape[-1]
        sigma, current = _host(conductivity), _host(currents)
        k = forward_operator._geometric_factors().detach()
        resistance = forward_operator.resistance_series(_steps(sigma, cells))
        rhoa = (k * resistance / current).reshape(*sigma.shape[:-1], -1)
        ctx.op = forward_operator
        ctx.sigma_meta = (conductivity.device, conductivity.dtype)
        ctx.current_meta = (currents.device, currents.dtype, tuple(currents.shape))
        ctx.save_for_backward(sigma, current, rhoa, k)
        ctx.save_for_forward(sigma, current, rhoa, k)
        return rhoa.to(device=conductivity.device, dtype=conductivity.dtype)

    @staticmethod
    @once_differentiable
    def backward(ctx: Any, grad_output):
        sigma, current, rhoa, k = ctx.saved_tensors
        g = _host(grad_output)
        cells = sigma.shape[-1]
        grad_sigma = grad_current = None
        if ctx.needs_input_grad[0]:
            cotangents = _steps(g * k / current, rhoa.shape[-1])
            grad_sigma = ctx.op.vjp_series(_steps(sigma, cells), cotangents).reshape(sigma.shape)
            device, dtype = ctx.sigma_meta
            grad_sigma = grad_sigma.to(device=device, dtype=dtype)
        if ctx.needs_input_grad[1]:
            device, dtype, shape = ctx.current_meta
            per_datum = _steps(-g * rhoa / current, rhoa.shape[-1]).sum(dim=0)  # summed over steps
            grad_current = per_datum.sum().reshape(shape) if shape in ((), (1,)) else per_datum
            grad_current = grad_current.to(device=device, dtype=dtype)
        return grad_sigma, grad_current, None

    @staticmethod
    def jvp(ctx: Any, conductivity_tangent, currents_tangent, forward_operator_tangent):
        del forward_operator_tangent
        sigma, current, rhoa, k = ctx.saved_tensors
        cells = sigma.shape[-1]
        tangent = torch.zeros_like(rhoa)
        if conductivity_tangent is not None:
            models, directions = _steps(sigma, cells), _steps(_host(conductivity_tangent), cells)
            dR = torch.stack([ctx.op.jvp(m, d) for m, d in zip(models, directions)]).reshape(rhoa.shape)
            tangent += k * dR / current
        if currents_tangent is not None:
            tangent -= rhoa * _host(currents_tangent) / current
        device, dtype = ctx.sigma_meta
        return tangent.to(device=device, dtype=dtype)


ERT3DApparentResistivityFunction = ERT2p5DApparentResistivityFunction
ERTApparentResistivityFunction = ERT2p5DApparentResistivityFunction


def apparent_resistivity_autograd(conductivity: torch.Tensor, forward_operator, currents=1.0) -> torch.Tensor:
    """Differentiable apparent resistivity (``forward/autograd.py:188-207``); ``currents`` is
    converted to the conductivity's device and dtype."""

    if not isinstance(conductivity, torch.Tensor):
        raise TypeError("conductivity must be a torch.Tensor")
    current = torch.as_tensor(currents, device=conductivity.device, dtype=conductivity.dtype)
    return ERT2p5DApparentResistivityFunction.apply(conductivity, current, forward_operator)
