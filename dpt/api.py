"""``project()``: the direct real-volume to device-projection path.

Same call contract as python/dpt/api.py:18-98.  ``ImagingVolume`` here is the
minimal record that call needs (values in mm^-1 on a ``GridSpec``); file
readers and HU conversion are not part of this package.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .contracts import ContractError
from .volumes import GridSpec


@dataclass(frozen=True)
class ImagingVolume:
    """Host attenuation values (nz,ny,nx) on ``grid`` with declared units."""

    values: np.ndarray
    grid: GridSpec
    value_units: str = "mm^-1"

    def __post_init__(self) -> None:
        if not isinstance(self.grid, GridSpec):
            raise ContractError("grid must be a GridSpec")
        if tuple(np.shape(self.values)) != self.grid.shape:
            raise ContractError("values must have the grid shape (nz,ny,nx)")


def project(
    volume,
    geometry=None,
    *,
    device=None,
    precision=None,
    samples_per_ray=None,
    step_mm=None,
    pose=None,
    packed_pose=None,
    prepared=None,
    out=None,
):
    """Return device optical depth (views,rows,columns) from attenuation.

    New calls default to the first ROCm device, FP32 and 1024 midpoint samples.
    Pass ``prepared=Projector(...)`` to reuse geometry and device plans.
    """
    import torch

    from .torch import Projector
    from .torch.modules import _geometry

    if prepared is not None and not isinstance(prepared, Projector):
        raise ContractError("prepared must be a dpt.torch.Projector")
    if isinstance(volume, ImagingVolume):
        if volume.value_units != "mm^-1":
            raise ContractError("projection requires attenuation units='mm^-1'; convert HU first")
        grid = volume.grid
        target = "cuda" if device is None else device
        field = torch.tensor(np.asarray(volume.values), dtype=torch.float32, device=target)
    elif isinstance(volume, torch.Tensor):
        if prepared is None:
            raise ContractError("a device tensor requires prepared=Projector(grid=..., geometry=...)")
        grid = prepared.grid
        field = volume
        if device is not None:
            requested = torch.device(device)
            same = requested == field.device or (requested.type == field.device.type and requested.index is None)
            if not same:
                raise ContractError("device must match the supplied tensor; move it explicitly")
    else:
        raise ContractError("volume must be an ImagingVolume or a prepared device tensor")
    if prepared is None:
        if geometry is None:
            raise ContractError("geometry is required when no prepared Projector is supplied")
        prepared = Projector(
            grid=grid,
            geometry=geometry,
            precision="fp32" if precision is None else precision,
            samples_per_ray=samples_per_ray,
            step_mm=step_mm,
        )
    else:
        if grid != prepared.grid:
            raise ContractError("volume grid does not match the prepared Projector")
        if geometry is not None and _geometry(geometry) != prepared.geometry:
            raise ContractError("geometry does not match the prepared Projector")
        if precision is not None and precision != prepared.precision:
            raise ContractError("precision does not match the prepared Projector")
        if samples_per_ray is not None and samples_per_ray != prepared.samples_per_ray:
            raise ContractError("samples_per_ray does not match the prepared Projector")
        if step_mm is not None and step_mm != prepared.step_mm:
            raise ContractError("step_mm does not match the prepared Projector")
    return prepared(field, pose, packed_pose=packed_pose, out=out)
