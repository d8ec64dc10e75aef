"""Differentiable SE(3) chart ``G(xi) = anchor @ exp(hat(xi))`` in Torch.

Restates python/dpt/torch/geometry.py:19-63: packed poses are twelve float64
values (row-major rotation, then translation in mm); ``xi`` is
(tx,ty,tz,rx,ry,rz) in mm and radians, acting on the right (object axes).

Device tensors take ``dpt::compose_pose``: one HIP launch forms the chart
and the exact partial derivatives of the computed exponential, one more
gives the gradients (``dpt_compose_pose`` / ``dpt_compose_pose_vjp``);
``torch.linalg.matrix_exp`` of the same twist is kept for host tensors.
The two agree to rounding.
"""

from __future__ import annotations

import torch

from .. import _native
from ..contracts import ContractError


@torch.library.custom_op("dpt::compose_pose", mutates_args=())
def _compose_device(anchor: torch.Tensor, tangent: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """(count,12) anchors and (count,6) tangents, contiguous float64 on one
    device -> (count,12) composed poses and the (count,84) jet."""
    count = anchor.shape[0]
    out = torch.empty((count, 12), dtype=torch.float64, device=anchor.device)
    jet = torch.empty((count, _native.POSE_JET), dtype=torch.float64, device=anchor.device)
    with torch.cuda.device(anchor.device):
        stream = torch.cuda.current_stream(anchor.device).cuda_stream
        _native.compose_pose(anchor.data_ptr(), tangent.data_ptr(), count, out.data_ptr(), jet.data_ptr(), stream)
    return out, jet


@_compose_device.register_fake
def _compose_device_fake(anchor, tangent):
    count = anchor.shape[0]
    return anchor.new_empty((count, 12)), anchor.new_empty((count, _native.POSE_JET))


@torch.library.custom_op("dpt::compose_pose_vjp", mutates_args=())
def _compose_device_vjp(anchor: torch.Tensor, jet: torch.Tensor, grad: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    count = anchor.shape[0]
    grad = grad.contiguous()
    grad_anchor = torch.empty((count, 12), dtype=torch.float64, device=anchor.device)
    grad_tangent = torch.empty((count, 6), dtype=torch.float64, device=anchor.device)
    with torch.cuda.device(anchor.device):
        stream = torch.cuda.current_stream(anchor.device).cuda_stream
        _native.compose_pose_vjp(anchor.data_ptr(), jet.data_ptr(), grad.data_ptr(), count, grad_anchor.data_ptr(),
                                 grad_tangent.data_ptr(), stream)
    return grad_anchor, grad_tangent


@_compose_device_vjp.register_fake
def _compose_device_vjp_fake(anchor, jet, grad):
    count = anchor.shape[0]
    return anchor.new_empty((count, 12)), anchor.new_empty((count, 6))


def _compose_setup(ctx, inputs, output):
    anchor, _ = inputs
    _, jet = output
    ctx.save_for_backward(anchor, jet)
    ctx.mark_non_differentiable(jet)


def _compose_backward(ctx, grad_out, grad_jet):
    anchor, jet = ctx.saved_tensors
    if grad_out is None:
        return None, None
    grad_anchor, grad_tangent = _compose_device_vjp(anchor, jet, grad_out)
    return grad_anchor, grad_tangent


_compose_device.register_autograd(_compose_backward, setup_context=_compose_setup)


def compose_pose(anchor: torch.Tensor, tangent: torch.Tensor) -> torch.Tensor:
    if not isinstance(anchor, torch.Tensor) or not isinstance(tangent, torch.Tensor):
        raise ContractError("anchor and tangent must be Torch tensors")
    if anchor.dtype != torch.float64 or tangent.dtype != torch.float64:
        raise ContractError("rigid pose charts require float64 tensors")
    if anchor.device != tangent.device:
        raise ContractError("anchor and tangent must share a device")
    if anchor.ndim < 1 or anchor.shape[-1] != 12 or tangent.ndim < 1 or tangent.shape[-1] != 6:
        raise ContractError("anchor and tangent require trailing shapes (12,) and (6,)")
    if any(count == 0 for count in (*anchor.shape, *tangent.shape)):
        raise ContractError("pose charts require nonempty tensor dimensions")
    try:
        lead = torch.broadcast_shapes(anchor.shape[:-1], tangent.shape[:-1])
    except RuntimeError as error:
        raise ContractError("anchor and tangent leading dimensions do not broadcast") from error
    if anchor.device.type == "cuda":
        flat_anchor = anchor.expand(*lead, 12).reshape(-1, 12).contiguous()
        flat_tangent = tangent.expand(*lead, 6).reshape(-1, 6).contiguous()
        composed, _ = _compose_device(flat_anchor, flat_tangent)
        return composed.reshape(*lead, 12)
    rho, w = tangent[..., :3], tangent[..., 3:]
    wx, wy, wz = w.unbind(-1)
    zero = torch.zeros_like(wx)
    hat = torch.stack((zero, -wz, wy, wz, zero, -wx, -wy, wx, zero), dim=-1).reshape(*tangent.shape[:-1], 3, 3)
    upper = torch.cat((hat, rho.unsqueeze(-1)), dim=-1)
    lower = torch.zeros((*tangent.shape[:-1], 1, 4), dtype=tangent.dtype, device=tangent.device)
    increment = torch.linalg.matrix_exp(torch.cat((upper, lower), dim=-2))
    rotation = anchor[..., :9].reshape(*anchor.shape[:-1], 3, 3)
    translation = (rotation @ increment[..., :3, 3].unsqueeze(-1)).squeeze(-1) + anchor[..., 9:]
    composed = rotation @ increment[..., :3, :3]
    return torch.cat((composed.reshape(*lead, 9), translation.expand(*lead, 3)), dim=-1).contiguous()
