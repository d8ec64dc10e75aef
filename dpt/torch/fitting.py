"""A rigid-pose objective for :func:`dpt.registration.recover_parameters`.

The chart of the search is a scaled pose tangent at the Projector's anchor:
``xi = scales * chi`` (tx, ty, tz in mm, rx, ry, rz in radians), the
objective ``1/2 sum (projector(field, xi) - target)^2`` in FP64 with its
chart gradient from the projector's pose VJP (one forward and one pose VJP
on the device per evaluation; only scalars return to the host).  A point
whose rotation norm exceeds ``rotation_radius`` is a rejectable trial, as the
reference's pose charts reject points outside their radius
(``PoseChart``, exercised by tests/python/test_recovery_trial_diagnostics.py).
"""

from __future__ import annotations

import math

import torch

from ..contracts import ContractError, TrialDomainError
from ..registration import Evaluation, Vector


class PoseObjective:
    def __init__(self, projector, field: torch.Tensor, target: torch.Tensor, *, scales=(1.0,) * 6,
                 rotation_radius: float | None = None) -> None:
        scales = tuple(float(s) for s in scales)
        if len(scales) != 6 or not all(math.isfinite(s) and s > 0.0 for s in scales):
            raise ContractError("scales must be six positive finite chart units")
        if tuple(target.shape) != projector.image_shape or target.device != field.device:
            raise ContractError("target must match the projector's image shape on the field device")
        if rotation_radius is not None and not (math.isfinite(rotation_radius) and rotation_radius > 0.0):
            raise ContractError("rotation_radius must be positive and finite")
        self.projector, self.field = projector, field.detach()
        self.target = target.detach().double()
        self.scales = scales
        self.rotation_radius = rotation_radius

    def tangent(self, parameters: Vector) -> Vector:
        """Chart point -> pose tangent xi (mm, radians)."""
        if len(parameters) != 6:
            raise ContractError("a rigid-pose chart has six parameters")
        xi = tuple(s * float(p) for s, p in zip(self.scales, parameters))
        if self.rotation_radius is not None and math.hypot(*xi[3:]) > self.rotation_radius:
            raise TrialDomainError("rotation outside the chart radius")
        return xi

    def __call__(self, parameters: Vector) -> Evaluation:
        xi = self.tangent(parameters)
        pose = torch.tensor(xi, dtype=torch.float64, device=self.field.device, requires_grad=True)
        residual = self.projector(self.field, pose).double() - self.target
        loss = 0.5 * torch.dot(residual.reshape(-1), residual.reshape(-1))
        (gradient,) = torch.autograd.grad(loss, pose)
        values = gradient.tolist()
        return Evaluation(float(loss), tuple(g * s for g, s in zip(values, self.scales)))
