"""Sampled scalar-field metadata (python/dpt/volumes.py:17-93).

The sample buffer is contiguous (nz, ny, nx), x fastest, in mm^-1.  The grid
origin is the first sample centre; support extends half a spacing beyond the
outer centres, inside which interpolation clamps to the outer sample.
"""

from __future__ import annotations

from dataclasses import dataclass

from .contracts import ContractError
from .geometry import IDENTITY, Matrix3, Vector3, matvec, rotation_matrix, transpose, vector3


@dataclass(frozen=True)
class GridSpec:
    """Oriented anisotropic grid: shape (nz,ny,nx), spacing (x,y,z) in mm."""

    shape: tuple[int, int, int]
    spacing_mm: Vector3
    origin_mm: Vector3 = (0.0, 0.0, 0.0)
    orientation: Matrix3 = IDENTITY

    def __post_init__(self) -> None:
        if not isinstance(self.shape, tuple) or len(self.shape) != 3 or any(
            type(x) is not int or x < 1 for x in self.shape
        ):
            raise ContractError("grid shape is (nz, ny, nx), each a positive integer")
        object.__setattr__(self, "spacing_mm", vector3(self.spacing_mm, "spacing_mm"))
        object.__setattr__(self, "origin_mm", vector3(self.origin_mm, "origin_mm"))
        object.__setattr__(self, "orientation", rotation_matrix(self.orientation, "orientation"))
        if min(self.spacing_mm) <= 0:
            raise ContractError("grid spacing must be positive")
        if self.voxels > 2**31 - 1:
            raise ContractError("grid exceeds signed 32-bit indexing")

    @property
    def voxels(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]

    @property
    def xyz(self) -> tuple[int, int, int]:
        return (self.shape[2], self.shape[1], self.shape[0])

    def object_to_grid_matrix(self) -> tuple[float, ...]:
        """diag(1/h) O^T, row-major: G[i][j] = O[3j+i] / h[i] (projection.py:571-576)."""
        return tuple(
            self.orientation[3 * j + i] / self.spacing_mm[i] for i in range(3) for j in range(3)
        )

    def object_to_grid(self, point_mm) -> Vector3:
        d = tuple(point_mm[i] - self.origin_mm[i] for i in range(3))
        a = matvec(transpose(self.orientation), d)
        return (a[0] / self.spacing_mm[0], a[1] / self.spacing_mm[1], a[2] / self.spacing_mm[2])

    def grid_to_object(self, index) -> Vector3:
        s = tuple(index[i] * self.spacing_mm[i] for i in range(3))
        a = matvec(self.orientation, s)
        return (a[0] + self.origin_mm[0], a[1] + self.origin_mm[1], a[2] + self.origin_mm[2])

    @property
    def support(self) -> tuple[Vector3, Vector3]:
        nx, ny, nz = self.xyz
        return (-0.5, -0.5, -0.5), (nx - 0.5, ny - 0.5, nz - 0.5)

    def centre_mm(self) -> Vector3:
        nx, ny, nz = self.xyz
        return self.grid_to_object(((nx - 1) / 2, (ny - 1) / 2, (nz - 1) / 2))

    def support_corners_mm(self) -> list[Vector3]:
        nx, ny, nz = self.xyz
        return [
            self.grid_to_object((x, y, z))
            for x in (-0.5, nx - 0.5)
            for y in (-0.5, ny - 0.5)
            for z in (-0.5, nz - 0.5)
        ]

    def encode(self) -> dict:
        return {
            "shape": list(self.shape),
            "spacing_mm": list(self.spacing_mm),
            "origin_mm": list(self.origin_mm),
            "orientation": list(self.orientation),
        }

    @classmethod
    def decode(cls, value: dict) -> GridSpec:
        return cls(
            tuple(int(x) for x in value["shape"]),
            tuple(float(x) for x in value["spacing_mm"]),
            tuple(float(x) for x in value["origin_mm"]),
            tuple(float(x) for x in value["orientation"]),
        )
