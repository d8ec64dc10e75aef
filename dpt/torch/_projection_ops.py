"""Checked multiview Torch operators over the native HIP plan.

Same operator contract as the reference (python/dpt/torch/_projection_ops.py:
256-411), registered in the ``dpt`` namespace:

* ``dpt::project(field, poses, value, active_volume, active_pose) -> image``
* ``dpt::project_vjp(field, poses, cotangent, value, active_volume,
  active_pose) -> (grad_field | empty(0), grad_poses | empty(0))``
* ``dpt::project_out(field, poses, value, out!)``

plus ``adjoint_gather``, the deterministic FP64 transpose (an explicit,
non-autograd call).

``value`` is the JSON configuration.  Fakes make the ops traceable; autograd
is first order only.  Work is enqueued on the caller's current stream (one
native plan per stream), followed by exactly one status readback per call.
Device tensors only: there is no CPU execution path.
"""

from __future__ import annotations

import torch

from .. import _native
from ..contracts import ContractError
from ..projection import ProjectionSpec, configuration as _encode, consume, decode, prepare

_TAGS = (torch.Tag.cudagraph_unsafe,)


def configuration(grid, geometry, spec: ProjectionSpec) -> str:
    return _encode(grid, geometry, spec)


def _metadata(field: torch.Tensor, poses: torch.Tensor, value: str):
    decoded = decode(value)
    if tuple(field.shape) != decoded.grid.shape or field.dtype != torch.float32 or not field.is_contiguous():
        raise ContractError("field must use contiguous float32 storage with grid shape")
    if tuple(poses.shape) != (decoded.views, 12) or poses.dtype != torch.float64 or poses.device != field.device:
        raise ContractError("poses must be float64 (views,12) on the field device")
    dtype = torch.float64 if decoded.spec.precision == "float64" else torch.float32
    return decoded.image_shape, dtype


def _require_device(field: torch.Tensor) -> None:
    if field.device.type != "cuda":
        raise ContractError(
            "dpt executes on ROCm GPUs only; move field and poses to a cuda device"
        )


def _needs(active_volume: bool, active_pose: bool) -> int:
    return (_native.VOLUME if active_volume else 0) | (_native.POSE if active_pose else 0)


def _field_pointer(entry, field: torch.Tensor, stream: int) -> int:
    """Pointer to the field in the plan's storage dtype (converting if half)."""
    storage = entry.decoded.spec.storage
    if storage == "float32":
        return field.data_ptr()
    dtype = torch.bfloat16 if storage == "bfloat16" else torch.float16
    half = torch.empty(field.numel() + _native.HALF_PAD_ELEMENTS, dtype=dtype, device=field.device)
    _native.convert_field(field.data_ptr(), half.data_ptr(), field.numel(), storage, None, stream)
    entry.half = half  # keep alive until the next call on this plan
    return half.data_ptr()


def _forward(field, poses, value, active_volume, active_pose, out=None):
    shape, dtype = _metadata(field, poses, value)
    if out is not None and (
        tuple(out.shape) != shape or out.dtype != dtype or out.device != field.device or not out.is_contiguous()
    ):
        raise ContractError("out must match the projector's shape, dtype and input device")
    if out is not None and any(
        out.untyped_storage().data_ptr() == x.untyped_storage().data_ptr() for x in (field, poses)
    ):
        raise ContractError("out must not share storage with projection inputs")
    _require_device(field)
    stream = torch.cuda.current_stream(field.device).cuda_stream
    entry = prepare(value, field.device.index or 0, stream, _needs(active_volume, active_pose))
    with entry.lock, torch.no_grad():
        packed = poses.contiguous()
        images = torch.empty(shape, dtype=dtype, device=field.device) if out is None else out
        pointer = _field_pointer(entry, field, stream)
        entry.plan.forward(pointer, packed.data_ptr(), images.data_ptr(), stream)
        consume(entry, stream)
    return images


@torch.library.custom_op("dpt::project", mutates_args=(), tags=_TAGS)
def project(
    field: torch.Tensor, poses: torch.Tensor, value: str, active_volume: bool, active_pose: bool
) -> torch.Tensor:
    return _forward(field, poses, value, active_volume, active_pose)


@project.register_fake
def _project_fake(field, poses, value, active_volume, active_pose):
    shape, dtype = _metadata(field, poses, value)
    return torch.empty(shape, dtype=dtype, device=field.device)


@torch.library.custom_op("dpt::project_vjp", mutates_args=(), tags=_TAGS)
def project_vjp(
    field: torch.Tensor,
    poses: torch.Tensor,
    cotangent: torch.Tensor,
    value: str,
    active_volume: bool,
    active_pose: bool,
) -> tuple[torch.Tensor, torch.Tensor]:
    shape, dtype = _metadata(field, poses, value)
    if tuple(cotangent.shape) != shape or cotangent.dtype != dtype or cotangent.device != field.device:
        raise ContractError("cotangent must match the projected image shape, dtype and device")
    if not active_volume and not active_pose:
        raise ContractError("request at least one projection gradient")
    _require_device(field)
    stream = torch.cuda.current_stream(field.device).cuda_stream
    entry = prepare(value, field.device.index or 0, stream, _needs(active_volume, active_pose))
    with entry.lock, torch.no_grad():
        packed = poses.contiguous()
        seed = cotangent.contiguous()
        gradient = torch.empty_like(field) if active_volume else field.new_empty((0,))
        pose_gradient = torch.empty_like(packed) if active_pose else poses.new_empty((0,))
        flags = _needs(active_volume, active_pose)
        pointer = _field_pointer(entry, field, stream)
        entry.plan.vjp(
            pointer,
            packed.data_ptr(),
            seed.data_ptr(),
            gradient.data_ptr() if active_volume else None,
            pose_gradient.data_ptr() if active_pose else None,
            flags,
            stream,
        )
        consume(entry, stream)
    return gradient, pose_gradient


@project_vjp.register_fake
def _vjp_fake(field, poses, cotangent, value, active_volume, active_pose):
    shape, dtype = _metadata(field, poses, value)
    if tuple(cotangent.shape) != shape or cotangent.dtype != dtype or cotangent.device != field.device:
        raise ContractError("cotangent must match the projected image shape, dtype and device")
    return (
        field.new_empty(field.shape if active_volume else (0,)),
        poses.new_empty(poses.shape if active_pose else (0,)),
    )


def _setup_context(ctx, inputs, output):
    field, poses, value, _, _ = inputs
    ctx.save_for_backward(field, poses)
    ctx.configuration = value


def _backward(ctx, cotangent):
    if torch.is_grad_enabled():
        raise ContractError("the Torch projection adapter supports first-order derivatives only")
    field, poses = ctx.saved_tensors
    active_volume, active_pose = ctx.needs_input_grad[:2]
    gradient, pose_gradient = project_vjp(field, poses, cotangent, ctx.configuration, active_volume, active_pose)
    return (
        gradient if active_volume else None,
        pose_gradient if active_pose else None,
        None,
        None,
        None,
    )


project.register_autograd(_backward, setup_context=_setup_context)


def ray_vjp(field: torch.Tensor, poses: torch.Tensor, cotangent: torch.Tensor, value: str) -> torch.Tensor:
    """Per ray: cotangent x d(optical depth)/d(world source, world endpoint),
    shape (views, rows, columns, 6) float64 (projection_ray_vjp,
    python/dpt/batched_projection.py:37-111).  Field, pose and grid are held
    fixed; the plan's arithmetic and integration rule are used."""
    shape, dtype = _metadata(field, poses, value)
    if tuple(cotangent.shape) != shape or cotangent.dtype != dtype or cotangent.device != field.device:
        raise ContractError("cotangent must match the projected image shape, dtype and device")
    _require_device(field)
    stream = torch.cuda.current_stream(field.device).cuda_stream
    entry = prepare(value, field.device.index or 0, stream, 0)
    with entry.lock, torch.no_grad():
        packed = poses.contiguous()
        seed = cotangent.contiguous()
        out = torch.empty((*shape, 6), dtype=torch.float64, device=field.device)
        pointer = _field_pointer(entry, field, stream)
        entry.plan.ray_vjp(pointer, packed.data_ptr(), seed.data_ptr(), out.data_ptr(), 0, stream)
        consume(entry, stream)
    return out


def pose_sensitivities(field: torch.Tensor, poses: torch.Tensor, cotangent: torch.Tensor, value: str) -> torch.Tensor:
    """Per ray: cotangent x d(optical depth)/d(delta) for pose @ exp(delta^),
    shape (views, rows, columns, 6) float64, rows ordered tx, ty, tz, rx, ry,
    rz (projection_pose_sensitivities, python/dpt/projection.py:1021-1084)."""
    shape, dtype = _metadata(field, poses, value)
    if tuple(cotangent.shape) != shape or cotangent.dtype != dtype or cotangent.device != field.device:
        raise ContractError("cotangent must match the projected image shape, dtype and device")
    _require_device(field)
    stream = torch.cuda.current_stream(field.device).cuda_stream
    entry = prepare(value, field.device.index or 0, stream, 0)
    with entry.lock, torch.no_grad():
        packed = poses.contiguous()
        seed = cotangent.contiguous()
        out = torch.empty((*shape, 6), dtype=torch.float64, device=field.device)
        pointer = _field_pointer(entry, field, stream)
        entry.plan.pose_sensitivities(pointer, packed.data_ptr(), seed.data_ptr(), out.data_ptr(), 0, stream)
        consume(entry, stream)
    return out


def reduce_rays(rays: torch.Tensor, jacobians: torch.Tensor, value: str) -> torch.Tensor:
    """out[v, k] = sum_p rays[v, p, :] . jacobians[v, k, :] in a fixed order
    (the intrinsic reduction of batched_projection_vjp,
    kernels/batched_projection.py:15-89)."""
    decoded = decode(value)
    views = decoded.views
    if tuple(rays.shape) != (*decoded.image_shape, 6) or rays.dtype != torch.float64:
        raise ContractError("rays must be float64 (views, rows, columns, 6)")
    if jacobians.dim() != 3 or jacobians.shape[0] != views or jacobians.shape[2] != 6:
        raise ContractError("jacobians must be float64 (views, components, 6)")
    _require_device(rays)
    stream = torch.cuda.current_stream(rays.device).cuda_stream
    entry = prepare(value, rays.device.index or 0, stream, 0)
    jac = jacobians.to(device=rays.device, dtype=torch.float64).contiguous()
    out = torch.empty((views, jac.shape[1]), dtype=torch.float64, device=rays.device)
    with entry.lock:
        entry.plan.reduce_rays(rays.contiguous().data_ptr(), jac.data_ptr(), jac.shape[1], out.data_ptr(), 0, stream)
        consume(entry, stream)
    return out


def adjoint_gather(poses: torch.Tensor, cotangent: torch.Tensor, value: str, out: torch.Tensor,
                   accumulate: bool) -> torch.Tensor:
    """Deterministic FP64 transpose: out[v] (+)= sum_p A(poses)[p, v] cotangent[p].

    Same contract as projection_adjoint_gather (python/dpt/projection.py:
    1253-1405) for a multiview plan: strict float64 sampling, a float64
    cotangent of the image shape, out float64 or float32 with the grid's voxel
    count, accumulate adding the previous out after each voxel's FP64 sum.
    Invalid input raises before anything is written.
    """
    decoded = decode(value)
    if decoded.spec.precision != "float64":
        raise ContractError("deterministic gather requires strict float64 sampling")
    if type(accumulate) is not bool:
        raise ContractError("accumulate must be a boolean")
    device = cotangent.device
    if tuple(cotangent.shape) != decoded.image_shape or cotangent.dtype != torch.float64:
        raise ContractError("cotangent must be float64 with the projected image shape")
    if tuple(poses.shape) != (decoded.views, 12) or poses.dtype != torch.float64 or poses.device != device:
        raise ContractError("poses must be float64 (views,12) on the cotangent device")
    if out.dtype not in (torch.float32, torch.float64) or out.numel() != decoded.grid.voxels:
        raise ContractError("out must be float32 or float64 with one element per voxel")
    if out.device != device or not out.is_contiguous():
        raise ContractError("out must be contiguous on the cotangent device")
    if any(out.untyped_storage().data_ptr() == x.untyped_storage().data_ptr() for x in (poses, cotangent)):
        raise ContractError("out must not share storage with the gather inputs")
    _require_device(cotangent)
    stream = torch.cuda.current_stream(device).cuda_stream
    entry = prepare(value, device.index or 0, stream, 0)
    with entry.lock, torch.no_grad():
        packed = poses.contiguous()
        seed = cotangent.contiguous()
        precision = "float64" if out.dtype == torch.float64 else "float32"
        flags = _native.ACCUMULATE if accumulate else 0
        entry.plan.adjoint_gather(packed.data_ptr(), seed.data_ptr(), out.data_ptr(), precision, flags, stream)
        consume(entry, stream)
    return out


@torch.library.custom_op("dpt::project_out", mutates_args={"out"}, tags=_TAGS)
def project_out(field: torch.Tensor, poses: torch.Tensor, value: str, out: torch.Tensor) -> None:
    _forward(field, poses, value, False, False, out)


@project_out.register_fake
def _out_fake(field, poses, value, out):
    shape, dtype = _metadata(field, poses, value)
    if tuple(out.shape) != shape or out.dtype != dtype or out.device != field.device or not out.is_contiguous():
        raise ContractError("out must match the projector's shape, dtype and input device")
