"""PyTorch autograd bridge for the 2.5D and 3D operators.

Same contract as ``adtlert/forward/autograd.py:17-207``: forward returns apparent resistivity
of one model ``(n_cells,)`` or a series ``(n_steps, n_cells)``; reverse mode calls
``vjp_series`` (hand-written adjoint kernels), forward mode calls ``jvp``; gradients flow to
conductivity and to scalar / per-datum currents; second derivatives are not supported.
"""

from __future__ import annotations

from typing import Any

import torch
from torch.autograd.function import once_differentiable

from .dtypes import FLOAT_DTYPE
from .ert2p5d import ERTForward2p5D
from .ert3d import ERTForward3D


def _validate_inputs(conductivity, currents, forward_operator) -> None:
    if not isinstance(forward_operator, (ERTForward2p5D, ERTForward3D)):
        raise TypeError("forward_operator must be an ERTForward2p5D or ERTForward3D instance")
    if not isinstance(conductivity, torch.Tensor) or not conductivity.is_floating_point():
        raise TypeError("conductivity must be a floating-point torch.Tensor")
    cells = forward_operator.mesh.cell_count
    if conductivity.ndim not in (1, 2) or conductivity.shape[-1] != cells:
        raise ValueError(f"conductivity must have shape ({cells},) or (n_steps, {cells}), got {tuple(conductivity.shape)}")
    if not bool(torch.all(torch.isfinite(conductivity))) or bool(torch.any(conductivity <= 0.0)):
        raise ValueError("conductivity must contain finite positive values")
    if not isinstance(currents, torch.Tensor) or not currents.is_floating_point():
        raise TypeError("currents must be a floating-point torch.Tensor")
    count = forward_operator.survey.measurement_count
    if tuple(currents.shape) not in {(), (1,), (count,)}:
        raise ValueError(f"currents must be scalar or have shape ({count},), got {tuple(currents.shape)}")
    if not bool(torch.all(torch.isfinite(currents))) or bool(torch.any(currents == 0.0)):
        raise ValueError("currents must contain finite non-zero values")


def _reduce_current_gradient(gradient: torch.Tensor, shape: torch.Size) -> torch.Tensor:
    if len(shape) == 0:
        return gradient.sum()
    if tuple(shape) == (1,):
        return gradient.sum().reshape(shape)
    return gradient


class ERT2p5DApparentResistivityFunction(torch.autograd.Function):
    """Matrix-free 2.5D/3D apparent resistivity as a differentiable torch operation."""

    @staticmethod
    def forward(ctx: Any, conductivity, currents, forward_operator):
        _validate_inputs(conductivity, currents, forward_operator)
        sigma = conductivity.detach().to(device="cpu", dtype=FLOAT_DTYPE)
        cur = currents.detach().to(device="cpu", dtype=FLOAT_DTYPE)
        scale = forward_operator._geometric_factors().detach()
        resistance = forward_operator.resistance_series(sigma.reshape(-1, sigma.shape[-1])).reshape(
            *sigma.shape[:-1], -1)
        rhoa = scale * resistance / cur
        ctx.forward_operator = forward_operator
        ctx.conductivity_device, ctx.conductivity_dtype = conductivity.device, conductivity.dtype
        ctx.currents_device, ctx.currents_dtype, ctx.currents_shape = currents.device, currents.dtype, currents.shape
        saved = (sigma, cur, rhoa, scale)
        ctx.save_for_backward(*saved)
        ctx.save_for_forward(*saved)
        return rhoa.to(device=conductivity.device, dtype=conductivity.dtype)

    @staticmethod
    @once_differentiable
    def backward(ctx: Any, grad_output):
        sigma, cur, rhoa, scale = ctx.saved_tensors
        g = grad_output.detach().to(device="cpu", dtype=FLOAT_DTYPE)
        grad_sigma = grad_cur = None
        if ctx.needs_input_grad[0]:
            cot = g * scale / cur
            grad_sigma = ctx.forward_operator.vjp_series(sigma.reshape(-1, sigma.shape[-1]),
                                                         cot.reshape(-1, cot.shape[-1])).reshape(sigma.shape)
            grad_sigma = grad_sigma.to(device=ctx.conductivity_device, dtype=ctx.conductivity_dtype)
        if ctx.needs_input_grad[1]:
            expanded = -g * rhoa / cur
            grad_cur = _reduce_current_gradient(expanded.reshape(-1, expanded.shape[-1]).sum(0), ctx.currents_shape)
            grad_cur = grad_cur.to(device=ctx.currents_device, dtype=ctx.currents_dtype)
        return grad_sigma, grad_cur, None

    @staticmethod
    def jvp(ctx: Any, conductivity_tangent, currents_tangent, forward_operator_tangent):
        del forward_operator_tangent
        sigma, cur, rhoa, scale = ctx.saved_tensors
        tangent = torch.zeros_like(rhoa)
        if conductivity_tangent is not None:
            direction = conductivity_tangent.detach().to(device="cpu", dtype=FLOAT_DTYPE)
            rt = torch.stack([ctx.forward_operator.jvp(m, d) for m, d in
                              zip(sigma.reshape(-1, sigma.shape[-1]), direction.reshape(-1, sigma.shape[-1]))])
            tangent = tangent + scale * rt.reshape(rhoa.shape) / cur
        if currents_tangent is not None:
            tangent = tangent - rhoa * currents_tangent.detach().to(device="cpu", dtype=FLOAT_DTYPE) / cur
        return tangent.to(device=ctx.conductivity_device, dtype=ctx.conductivity_dtype)


ERT3DApparentResistivityFunction = ERT2p5DApparentResistivityFunction
ERTApparentResistivityFunction = ERT2p5DApparentResistivityFunction


def apparent_resistivity_autograd(conductivity: torch.Tensor, forward_operator, currents=1.0) -> torch.Tensor:
    """Differentiable apparent resistivity (``forward/autograd.py:188-207``)."""

    if not isinstance(conductivity, torch.Tensor):
        raise TypeError("conductivity must be a torch.Tensor")
    current = torch.as_tensor(currents, device=conductivity.device, dtype=conductivity.dtype)
    return ERT2p5DApparentResistivityFunction.apply(conductivity, current, forward_operator)
