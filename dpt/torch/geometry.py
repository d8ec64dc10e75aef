"""Differentiable SE(3) chart ``G(xi) = anchor @ exp(hat(xi))`` in Torch.

Restates python/dpt/torch/geometry.py:19-63: packed poses are twelve float64
values (row-major rotation, then translation in mm); ``xi`` is
(tx,ty,tz,rx,ry,rz) in mm and radians, acting on the right (object axes).
"""

from __future__ import annotations

import torch

from ..contracts import ContractError


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
