"""``Projector``: prepared multiview projection as an ordinary PyTorch module.

Same constructor and call contract as the reference module
(python/dpt/torch/modules.py:69-228): a fixed millimetre grid and geometry,
``precision`` "fp32" (FP32 sampling and output) or "fp64", a fixed
``samples_per_ray`` (1024 by default) or one resolved from ``step_mm``, an
optional rigid ``anchor``; ``forward(field, pose=None, *, packed_pose=None,
out=None)`` returns optical depth (views, rows, columns).

Extensions: ``storage`` ("float32", "bfloat16", "float16") selects how the
field is held for sampling.  Half storage rounds the FP32 field once per call
and samples it with FP32 arithmetic and compensated along-ray sums; gradients
are returned for the FP32 field.  It requires ``precision="fp32"``.
``integration="cell_gauss"`` (with ``precision="fp64"``) exposes the
reference's exact cell-Gauss rule (``ProjectionSpec.integration``,
python/dpt/projection.py:107-135) through the module; forward and both
adjoints follow it.
"""

from __future__ import annotations

import math
from collections.abc import Sequence

import torch

from ..contracts import ContractError, finite_scalar, integer
from ..geometry import PANEL_TYPES, BatchedGeometry, ConeBeamIntrinsics, RayBundleGeometry, RigidTransform
from ..projection import ProjectionSpec
from ..volumes import GridSpec
from . import _projection_ops
from .geometry import compose_pose

__all__ = ["Projector"]


def _projection_spec(precision: str, samples: int, storage: str, integration: str = "midpoint") -> ProjectionSpec:
    if precision not in ("fp32", "fp64"):
        raise ContractError("precision must be fp32 or fp64")
    arithmetic = "float32" if precision == "fp32" else "float64"
    return ProjectionSpec(samples, arithmetic, storage, integration)


def _geometry(geometry):
    if isinstance(geometry, (BatchedGeometry, RayBundleGeometry)):
        return geometry
    if isinstance(geometry, PANEL_TYPES):
        return BatchedGeometry((geometry,))
    if isinstance(geometry, Sequence):
        return BatchedGeometry(tuple(geometry))
    raise ContractError("geometry must be a DetectorGeometry, BatchedGeometry or sequence of panels")


def _samples(grid: GridSpec, samples_per_ray, step_mm) -> int:
    """python/dpt/torch/modules.py:47-66."""
    if step_mm is None:
        return integer(1024 if samples_per_ray is None else samples_per_ray, "samples_per_ray", minimum=1)
    if samples_per_ray is not None:
        raise ContractError("choose samples_per_ray or step_mm")
    step = finite_scalar(step_mm, "step_mm", minimum=0.0)
    if step == 0.0:
        raise ContractError("step_mm must be positive")
    diagonal = math.hypot(*(n * h for n, h in zip(reversed(grid.shape), grid.spacing_mm)))
    count = diagonal / step
    if not math.isfinite(count) or count > 2**31 - 1:
        raise ContractError("step_mm requires too many midpoint samples")
    return max(1, math.ceil(count))


class Projector(torch.nn.Module):
    """Project device attenuation (nz,ny,nx) mm^-1 to optical depth (V,R,C)."""

    anchor: torch.Tensor | None

    def __init__(
        self,
        *,
        grid: GridSpec,
        geometry,
        precision: str = "fp32",
        samples_per_ray: int | None = None,
        step_mm: float | None = None,
        anchor: RigidTransform | torch.Tensor | None = None,
        storage: str = "float32",
        integration: str = "midpoint",
    ) -> None:
        super().__init__()
        if not isinstance(grid, GridSpec):
            raise ContractError("grid must be a GridSpec")
        self.grid = grid
        self.geometry = _geometry(geometry)
        self.precision = precision
        self.storage = storage
        self.samples_per_ray = _samples(grid, samples_per_ray, step_mm)
        self.step_mm = step_mm
        self.integration = integration
        self._spec = _projection_spec(precision, self.samples_per_ray, storage, integration)
        self._configuration = _projection_ops.configuration(self.grid, self.geometry, self._spec)
        self._fixed_poses: dict[str, torch.Tensor] = {}
        if isinstance(anchor, torch.Tensor):
            self._check_pose(anchor, 12, "anchor")
            self._fixed_anchor = None
            if isinstance(anchor, torch.nn.Parameter):
                self.register_parameter("anchor", anchor)
            else:
                self.register_buffer("anchor", anchor)
        else:
            if anchor is not None and not isinstance(anchor, RigidTransform):
                raise ContractError("anchor must be a RigidTransform or FP64 packed Torch tensor")
            self._fixed_anchor = RigidTransform() if anchor is None else anchor
            self.register_buffer("anchor", None)

    @property
    def views(self) -> int:
        return 1 if isinstance(self.geometry, RayBundleGeometry) else len(self.geometry.views)

    @property
    def image_shape(self) -> tuple[int, int, int]:
        if isinstance(self.geometry, RayBundleGeometry):
            return (1, 1, self.geometry.rays)
        return self.geometry.shape

    def _check_pose(self, value, components: int, name: str) -> None:
        shapes = ((components,), (self.views, components))
        if not isinstance(value, torch.Tensor) or value.dtype != torch.float64:
            raise ContractError(f"{name} must be a float64 Torch tensor")
        if tuple(value.shape) not in shapes:
            raise ContractError(f"{name} must have shape {shapes[0]} or {shapes[1]}")
        if not value.is_contiguous():
            raise ContractError(f"{name} must be contiguous")

    def _poses(self, field, pose, packed_pose) -> torch.Tensor:
        if packed_pose is not None:
            if pose is not None:
                raise ContractError("choose pose or packed_pose")
            self._check_pose(packed_pose, 12, "packed_pose")
            result = packed_pose
        else:
            anchor = self.anchor
            if anchor is None:
                key = str(field.device)
                if key not in self._fixed_poses:
                    self._fixed_poses[key] = torch.tensor(
                        self._fixed_anchor.packed(), dtype=torch.float64, device=field.device
                    )
                anchor = self._fixed_poses[key]
            if anchor.device != field.device:
                raise ContractError("move the Projector anchor to the field device with .to(device)")
            self._check_pose(anchor, 12, "anchor")
            if pose is None:
                result = anchor
            else:
                self._check_pose(pose, 6, "pose")
                if pose.device != field.device:
                    raise ContractError("pose and field must share a device")
                result = compose_pose(anchor, pose)
        if result.device != field.device:
            raise ContractError("packed_pose and field must share a device")
        return result.expand(self.views, 12) if result.ndim == 1 else result

    def _check_output(self, out, field, poses) -> None:
        if not isinstance(out, torch.Tensor):
            raise ContractError("out must be a Torch tensor")
        dtype = torch.float64 if self.precision == "fp64" else torch.float32
        if tuple(out.shape) != self.image_shape or out.dtype != dtype or out.device != field.device:
            raise ContractError("out must match the projector's shape, dtype and input device")
        if not out.is_contiguous():
            raise ContractError("out must be contiguous")
        if out.requires_grad or (torch.is_grad_enabled() and (field.requires_grad or poses.requires_grad)):
            raise ContractError("out is an inference buffer; omit it for autograd or use torch.no_grad()")

    def forward(self, field: torch.Tensor, pose: torch.Tensor | None = None, *, packed_pose=None, out=None):
        """Optical depth with field and optional rigid-pose gradients."""
        if not isinstance(field, torch.Tensor) or tuple(field.shape) != self.grid.shape:
            raise ContractError("field must be a Torch tensor with grid shape (nz,ny,nx)")
        if field.dtype != torch.float32 or not field.is_contiguous():
            raise ContractError("field must use contiguous float32 storage")
        if field.device.type not in ("cuda", "cpu", "meta"):
            raise ContractError("field must be on a ROCm (cuda) device")
        poses = self._poses(field, pose, packed_pose)
        if out is not None:
            self._check_output(out, field, poses)
            _projection_ops.project_out(field, poses, self._configuration, out)
            return out
        return _projection_ops.project(
            field,
            poses,
            self._configuration,
            field.requires_grad and torch.is_grad_enabled(),
            poses.requires_grad and torch.is_grad_enabled(),
        )

    def ray_vjp(self, field: torch.Tensor, cotangent: torch.Tensor, pose: torch.Tensor | None = None, *,
                packed_pose=None) -> torch.Tensor:
        """Per ray: cotangent x d(optical depth)/d(world source xyz, world
        endpoint xyz), shape (views, rows, columns, 6) float64, with the field
        and object pose fixed (projection_ray_vjp,
        python/dpt/batched_projection.py:37-111)."""
        if not isinstance(field, torch.Tensor) or tuple(field.shape) != self.grid.shape:
            raise ContractError("field must be a Torch tensor with grid shape (nz,ny,nx)")
        poses = self._poses(field, pose, packed_pose).detach()
        return _projection_ops.ray_vjp(field.detach(), poses.contiguous(), cotangent, self._configuration)

    def pose_sensitivities(self, field: torch.Tensor, cotangent: torch.Tensor | None = None,
                           pose: torch.Tensor | None = None, *, packed_pose=None) -> torch.Tensor:
        """Per ray: cotangent x d(optical depth)/d(delta) for G @ exp(delta^),
        G the packed pose this module evaluates, shape (views, rows, columns,
        6) float64 ordered tx, ty, tz, rx, ry, rz in mm and rad; without a
        cotangent the rows are the unweighted Jacobian.  Missed rays give zero
        rows (projection_pose_sensitivities,
        python/dpt/projection.py:1021-1084)."""
        if not isinstance(field, torch.Tensor) or tuple(field.shape) != self.grid.shape:
            raise ContractError("field must be a Torch tensor with grid shape (nz,ny,nx)")
        poses = self._poses(field, pose, packed_pose).detach()
        if cotangent is None:
            dtype = torch.float32 if self.precision == "fp32" else torch.float64
            cotangent = torch.ones(self.image_shape, dtype=dtype, device=field.device)
        return _projection_ops.pose_sensitivities(field.detach(), poses.contiguous(), cotangent, self._configuration)

    def intrinsics_vjp(self, field: torch.Tensor, cotangent: torch.Tensor, intrinsics, angles_rad,
                       pose: torch.Tensor | None = None, *, packed_pose=None) -> torch.Tensor:
        """Cone intrinsic cotangent (views, 4) = SOD, SDD, column offset, row
        offset per view (batched_projection_vjp with active intrinsics,
        python/dpt/batched_projection.py:340-539); sum the rows for a shared
        calibration.  The module's geometry must be ``intrinsics`` viewed at
        ``angles_rad`` (one record, or one per angle)."""
        angles = [float(a) for a in angles_rad]
        records = list(intrinsics) if isinstance(intrinsics, (list, tuple)) else [intrinsics] * len(angles)
        if len(records) != len(angles) or any(not isinstance(r, ConeBeamIntrinsics) for r in records):
            raise ContractError("intrinsics must be one cone record or one per angle")
        expected = BatchedGeometry(tuple(r.view(a) for r, a in zip(records, angles)))
        if expected != self.geometry:
            raise ContractError("the module's geometry is not these intrinsics at these angles")
        rays = self.ray_vjp(field, cotangent, pose, packed_pose=packed_pose)
        jac = torch.tensor(
            [[[*ds, *de] for ds, de in r.ray_derivatives(a)] for r, a in zip(records, angles)],
            dtype=torch.float64, device=rays.device,
        )
        return _projection_ops.reduce_rays(rays, jac, self._configuration)

    def adjoint_gather(
        self,
        cotangent: torch.Tensor,
        pose: torch.Tensor | None = None,
        *,
        packed_pose=None,
        out: torch.Tensor | None = None,
        volume_gradient_precision: str = "float64",
        accumulate: bool = False,
    ) -> torch.Tensor:
        """Deterministic volume adjoint: out[v] = sum_p A(pose)[p, v] c[p],
        each voxel summed in FP64 in a fixed (view, pixel, sample, corner)
        order, so repeated calls agree bit for bit.  Requires
        ``precision="fp64"``; ``cotangent`` is float64 (views, rows, columns);
        the result has the grid shape (nz, ny, nx) in
        ``volume_gradient_precision`` ("float64" or "float32").  With
        ``accumulate`` the sums are added to ``out``.  Same contract as
        projection_adjoint_gather (python/dpt/projection.py:1253-1405)."""
        if self.precision != "fp64":
            raise ContractError("deterministic gather requires strict float64 sampling (precision='fp64')")
        if volume_gradient_precision not in ("float32", "float64"):
            raise ContractError("volume_gradient_precision must be float32 or float64")
        if not isinstance(cotangent, torch.Tensor):
            raise ContractError("cotangent must be a Torch tensor")
        poses = self._poses(cotangent, pose, packed_pose)
        dtype = torch.float64 if volume_gradient_precision == "float64" else torch.float32
        if out is None:
            if accumulate:
                raise ContractError("accumulate needs an out tensor")
            out = torch.empty(self.grid.shape, dtype=dtype, device=cotangent.device)
        elif out.dtype != dtype:
            raise ContractError("out dtype must match volume_gradient_precision")
        return _projection_ops.adjoint_gather(poses.contiguous(), cotangent, self._configuration, out, accumulate)
