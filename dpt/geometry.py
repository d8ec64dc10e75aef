"""Right-handed rigid geometry in millimetres (host model of the projector).

Mirrors the reference's immutable geometry records and conventions
(python/dpt/geometry.py):

* rotations are row-major 3x3 tuples acting on column vectors;
* ``RigidTransform.packed()`` is twelve float64 values, rotation then
  translation (geometry.py:96-98);
* a flat ``DetectorGeometry`` has shape (rows, columns) and spacing
  (column pitch, row pitch); its ``origin_mm`` is the centre of pixel (0,0)
  (geometry.py:100-222);
* ``ConeBeamIntrinsics`` places the source at (0,-SOD,0) and the panel centre
  at (offset_column, SDD-SOD, offset_row) before a rotation about world z
  (geometry.py:617-701);
* ``compose_pose(anchor, xi)`` is ``anchor @ exp(xi^)`` with
  xi = (tx,ty,tz,rx,ry,rz) acting on the right (geometry.py:742-778).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import cast

from .contracts import ContractError, finite_tuple

Vector3 = tuple[float, float, float]
Matrix3 = tuple[float, float, float, float, float, float, float, float, float]
IDENTITY: Matrix3 = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


def vector3(value, name: str) -> Vector3:
    return cast(Vector3, finite_tuple(tuple(value), name, minimum=None, length=3))


def dot(a, b) -> float:
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b) -> Vector3:
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def matvec(m, v) -> Vector3:
    return cast(Vector3, tuple(m[3 * i] * v[0] + m[3 * i + 1] * v[1] + m[3 * i + 2] * v[2] for i in range(3)))


def transpose(m) -> Matrix3:
    return cast(Matrix3, tuple(m[3 * j + i] for i in range(3) for j in range(3)))


def matmul(a, b) -> Matrix3:
    return cast(
        Matrix3,
        tuple(sum(a[3 * i + k] * b[3 * k + j] for k in range(3)) for i in range(3) for j in range(3)),
    )


def rotation_matrix(value, name: str = "rotation") -> Matrix3:
    """Finite, orthonormal to 1e-10, right-handed (geometry.py:55-63)."""
    value = cast(Matrix3, finite_tuple(tuple(value), name, minimum=None, length=9))
    product = matmul(transpose(value), value)
    if max(abs(a - b) for a, b in zip(product, IDENTITY)) > 1e-10:
        raise ContractError(f"{name} must be orthonormal to 1e-10")
    columns = [(value[i], value[i + 3], value[i + 6]) for i in range(3)]
    if dot(cross(columns[0], columns[1]), columns[2]) < 0.0:
        raise ContractError(f"{name} must be right-handed")
    return value


@dataclass(frozen=True)
class RigidTransform:
    """T_WO maps object points into world coordinates; translation in mm."""

    rotation: Matrix3 = IDENTITY
    translation_mm: Vector3 = (0.0, 0.0, 0.0)

    def __post_init__(self) -> None:
        object.__setattr__(self, "rotation", rotation_matrix(self.rotation))
        object.__setattr__(self, "translation_mm", vector3(self.translation_mm, "translation"))

    def point(self, p) -> Vector3:
        r = matvec(self.rotation, p)
        return (r[0] + self.translation_mm[0], r[1] + self.translation_mm[1], r[2] + self.translation_mm[2])

    def direction(self, d) -> Vector3:
        return matvec(self.rotation, d)

    def inverse(self) -> RigidTransform:
        rt = transpose(self.rotation)
        t = matvec(rt, self.translation_mm)
        return RigidTransform(rt, (-t[0], -t[1], -t[2]))

    def compose(self, right: RigidTransform) -> RigidTransform:
        """self @ right (right applied first)."""
        return RigidTransform(matmul(self.rotation, right.rotation), self.point(right.translation_mm))

    def packed(self) -> tuple[float, ...]:
        return (*self.rotation, *self.translation_mm)


@dataclass(frozen=True)
class DetectorGeometry:
    """Flat cone panel; u increases columns, v increases rows.

    ``shape`` is (rows, columns); ``spacing_mm`` is (column pitch, row pitch).
    """

    source_mm: Vector3
    origin_mm: Vector3
    u: Vector3
    v: Vector3
    spacing_mm: tuple[float, float]
    shape: tuple[int, int]

    def __post_init__(self) -> None:
        for name in ("source_mm", "origin_mm", "u", "v"):
            object.__setattr__(self, name, vector3(getattr(self, name), name))
        if abs(dot(self.u, self.u) - 1.0) > 1e-10 or abs(dot(self.v, self.v) - 1.0) > 1e-10:
            raise ContractError("detector basis vectors must have unit length")
        if abs(dot(self.u, self.v)) > 1e-10:
            raise ContractError("detector basis vectors must be perpendicular")
        spacing = finite_tuple(tuple(self.spacing_mm), "pixel spacing", positive=True, length=2)
        _panel_shape(self.shape)
        object.__setattr__(self, "shape", tuple(self.shape))
        object.__setattr__(self, "spacing_mm", spacing)
        offset = tuple(self.source_mm[i] - self.origin_mm[i] for i in range(3))
        if abs(dot(offset, cross(self.u, self.v))) <= 1e-12:
            raise ContractError("source must lie outside the detector plane")

    @property
    def pixels(self) -> int:
        return self.shape[0] * self.shape[1]

    def pixel_centre(self, row: int, column: int) -> Vector3:
        if type(row) is not int or type(column) is not int:
            raise ContractError("pixel indices must be integers")
        if not (0 <= row < self.shape[0] and 0 <= column < self.shape[1]):
            raise ContractError("pixel index outside detector")
        du, dv = self.spacing_mm
        return cast(
            Vector3,
            tuple(self.origin_mm[i] + column * du * self.u[i] + row * dv * self.v[i] for i in range(3)),
        )

    def record(self) -> tuple[tuple[float, ...], int]:
        """Panel record for device ray generation (kernels/batched_geometry.py:9-39): mode 0."""
        return (*self.source_mm, *self.origin_mm, *self.u, *self.v, *self.spacing_mm, 0.0, 0.0, 0.0, 0.0), 0

    def packed(self) -> tuple[float, ...]:
        """source, pixel (0,0) centre, column step u*du, row step v*dv (projection.py:594-598)."""
        du, dv = self.spacing_mm
        return (
            *self.source_mm,
            *self.origin_mm,
            *(x * du for x in self.u),
            *(x * dv for x in self.v),
        )


def _pixel_index(row, column, shape) -> None:
    if type(row) is not int or type(column) is not int:
        raise ContractError("pixel indices must be integers")
    if not (0 <= row < shape[0] and 0 <= column < shape[1]):
        raise ContractError("pixel index outside detector")


@dataclass(frozen=True)
class ParallelGeometry:
    """Finite parallel segments ``s[r,c] = p[r,c] - length * direction``
    (geometry.py:306-366): ``origin_mm`` is the endpoint of pixel (0,0), ``u``
    increases columns, ``v`` rows; ``shape`` is (rows, columns) and
    ``spacing_mm`` (column, row), all in millimetres."""

    origin_mm: Vector3
    direction: Vector3
    length_mm: float
    u: Vector3
    v: Vector3
    spacing_mm: tuple[float, float]
    shape: tuple[int, int]

    def __post_init__(self) -> None:
        direction = vector3(self.direction, "parallel direction")
        if abs(dot(direction, direction) - 1.0) > 1e-10:
            raise ContractError("parallel direction must have unit length")
        length = finite_tuple((self.length_mm,), "parallel length", positive=True)[0]
        origin = vector3(self.origin_mm, "parallel origin")
        source = tuple(origin[i] - length * direction[i] for i in range(3))
        panel = DetectorGeometry(source, origin, self.u, self.v, self.spacing_mm, self.shape)
        object.__setattr__(self, "origin_mm", origin)
        object.__setattr__(self, "direction", direction)
        object.__setattr__(self, "length_mm", length)
        for name in ("u", "v", "spacing_mm", "shape"):
            object.__setattr__(self, name, getattr(panel, name))

    @property
    def pixels(self) -> int:
        return self.shape[0] * self.shape[1]

    def pixel_centre(self, row: int, column: int) -> Vector3:
        _pixel_index(row, column, self.shape)
        du, dv = self.spacing_mm
        return cast(
            Vector3,
            tuple(self.origin_mm[i] + column * du * self.u[i] + row * dv * self.v[i] for i in range(3)),
        )

    def segment(self, row: int, column: int) -> tuple[Vector3, Vector3]:
        """Source and endpoint of pixel (row, column)."""
        end = self.pixel_centre(row, column)
        return cast(Vector3, tuple(end[i] - self.length_mm * self.direction[i] for i in range(3))), end

    def record(self) -> tuple[tuple[float, ...], int]:
        return (*self.direction, *self.origin_mm, *self.u, *self.v, *self.spacing_mm, self.length_mm, 0.0, 0.0,
                0.0), 1


@dataclass(frozen=True)
class CurvedDetectorGeometry:
    """Cylindrical pixels ``p = axis + radius cos(theta) + u R sin(theta)``
    (geometry.py:369-431): the cylinder axis passes through ``source_mm``
    along ``v``; pixel (0,0) is ``origin_mm``; columns follow the tangent
    ``u`` with arc length ``spacing_mm[0]``, rows move along ``v`` by
    ``spacing_mm[1]``."""

    source_mm: Vector3
    origin_mm: Vector3
    u: Vector3
    v: Vector3
    spacing_mm: tuple[float, float]
    shape: tuple[int, int]

    def __post_init__(self) -> None:
        panel = DetectorGeometry(self.source_mm, self.origin_mm, self.u, self.v, self.spacing_mm, self.shape)
        for name in ("source_mm", "origin_mm", "u", "v", "spacing_mm", "shape"):
            object.__setattr__(self, name, getattr(panel, name))
        radial = self.radial_mm
        if not math.isfinite(self.radius_mm) or self.radius_mm <= 0.0:
            raise ContractError("curved detector radius must be finite and positive")
        if abs(dot(radial, self.u)) > 1e-10 * self.radius_mm:
            raise ContractError("curved detector u must be tangent to its source-centred cylinder")
        if self.spacing_mm[0] * (self.shape[1] - 1) >= 2 * math.pi * self.radius_mm:
            raise ContractError("curved detector columns must cover less than one full revolution")

    @property
    def radial_mm(self) -> Vector3:
        offset = tuple(self.origin_mm[i] - self.source_mm[i] for i in range(3))
        along = dot(offset, self.v)
        return cast(Vector3, tuple(offset[i] - along * self.v[i] for i in range(3)))

    @property
    def radius_mm(self) -> float:
        radial = self.radial_mm
        return math.sqrt(dot(radial, radial))

    @property
    def pixels(self) -> int:
        return self.shape[0] * self.shape[1]

    def pixel_centre(self, row: int, column: int) -> Vector3:
        _pixel_index(row, column, self.shape)
        radial, radius = self.radial_mm, self.radius_mm
        angle = column * self.spacing_mm[0] / radius
        return cast(
            Vector3,
            tuple(
                self.origin_mm[i]
                + (math.cos(angle) - 1.0) * radial[i]
                + math.sin(angle) * radius * self.u[i]
                + row * self.spacing_mm[1] * self.v[i]
                for i in range(3)
            ),
        )

    def record(self) -> tuple[tuple[float, ...], int]:
        return (*self.source_mm, *self.origin_mm, *self.u, *self.v, *self.spacing_mm, self.radius_mm,
                *self.radial_mm), 2


PANEL_TYPES = (DetectorGeometry, ParallelGeometry, CurvedDetectorGeometry)


@dataclass(frozen=True)
class RayBundleGeometry:
    """Explicit finite world segments, one output per segment (geometry.py:225-300)."""

    sources_mm: tuple[Vector3, ...]
    endpoints_mm: tuple[Vector3, ...]

    def __post_init__(self) -> None:
        if (
            not isinstance(self.sources_mm, tuple)
            or not isinstance(self.endpoints_mm, tuple)
            or not self.sources_mm
            or len(self.sources_mm) != len(self.endpoints_mm)
        ):
            raise ContractError("ray sources/endpoints must be equal non-empty immutable tuples")
        sources = tuple(vector3(s, "ray source") for s in self.sources_mm)
        endpoints = tuple(vector3(e, "ray endpoint") for e in self.endpoints_mm)
        if any(not math.isfinite(math.dist(s, e)) or s == e for s, e in zip(sources, endpoints)):
            raise ContractError("ray segments must have finite positive length")
        object.__setattr__(self, "sources_mm", sources)
        object.__setattr__(self, "endpoints_mm", endpoints)

    @property
    def rays(self) -> int:
        return len(self.sources_mm)

    @property
    def pixels(self) -> int:
        return self.rays

    def segment(self, ray: int) -> tuple[Vector3, Vector3]:
        return self.sources_mm[ray], self.endpoints_mm[ray]


@dataclass(frozen=True)
class BatchedGeometry:
    """O(views) panel metadata for an image of shape (views, rows, columns);
    flat cone, parallel and cylindrical panels may share one batch
    (geometry.py:462-520)."""

    views: tuple[DetectorGeometry | ParallelGeometry | CurvedDetectorGeometry, ...]

    def __post_init__(self) -> None:
        if (
            not isinstance(self.views, tuple)
            or not self.views
            or any(not isinstance(view, PANEL_TYPES) for view in self.views)
        ):
            raise ContractError("views must be a non-empty immutable tuple of panel geometries")
        if any(view.shape != self.views[0].shape for view in self.views):
            raise ContractError("batched views must share (rows,columns)")
        if self.pixels > 2**31 - 1:
            raise ContractError("batched geometry exceeds signed 32-bit indexing")

    @property
    def shape_rows_columns(self) -> tuple[int, int]:
        return self.views[0].shape

    @property
    def shape(self) -> tuple[int, int, int]:
        return (len(self.views), *self.views[0].shape)

    @property
    def pixels(self) -> int:
        return len(self.views) * self.views[0].pixels

    @property
    def flat_cone(self) -> bool:
        """Every view is a flat cone panel (the projected-window kernels apply)."""
        return all(type(view) is DetectorGeometry for view in self.views)

    def segment(self, view: int, row: int, column: int) -> tuple[Vector3, Vector3]:
        """One finite world segment (geometry.py:503-511)."""
        if type(view) is not int or not 0 <= view < len(self.views):
            raise ContractError("view index must be inside the batch")
        panel = self.views[view]
        if isinstance(panel, ParallelGeometry):
            return panel.segment(row, column)
        return panel.source_mm, panel.pixel_centre(row, column)


def _panel_shape(shape) -> None:
    if (
        not isinstance(shape, tuple)
        or len(shape) != 2
        or any(type(count) is not int or count <= 0 for count in shape)
    ):
        raise ContractError("detector shape must be (rows,columns), both positive integers")
    if shape[0] * shape[1] > 2**31 - 1:
        raise ContractError("detector exceeds signed 32-bit indexing")


@dataclass(frozen=True)
class ConeBeamIntrinsics:
    """Circular cone geometry with source (0,-SOD,0) before rotation about z."""

    source_object_distance_mm: float
    source_detector_distance_mm: float
    shape: tuple[int, int]
    spacing_mm: tuple[float, float]
    detector_offset_mm: tuple[float, float] = (0.0, 0.0)
    centre_mm: Vector3 = (0.0, 0.0, 0.0)

    def __post_init__(self) -> None:
        _panel_shape(self.shape)
        sod, sdd = finite_tuple(
            (self.source_object_distance_mm, self.source_detector_distance_mm),
            "SOD/SDD",
            positive=True,
            length=2,
        )
        if sdd <= sod:
            raise ContractError("source-detector distance must exceed source-object distance")
        object.__setattr__(self, "source_object_distance_mm", sod)
        object.__setattr__(self, "source_detector_distance_mm", sdd)
        object.__setattr__(
            self, "spacing_mm", finite_tuple(tuple(self.spacing_mm), "pixel spacing", positive=True, length=2)
        )
        object.__setattr__(
            self,
            "detector_offset_mm",
            finite_tuple(tuple(self.detector_offset_mm), "detector offset", minimum=None, length=2),
        )
        object.__setattr__(self, "centre_mm", vector3(self.centre_mm, "centre_mm"))

    def view(self, angle_rad: float) -> DetectorGeometry:
        """One view rotated by ``angle_rad`` about world z (geometry.py:662-682)."""
        angle = finite_tuple((float(angle_rad),), "view angle", minimum=None)[0]
        c, s = math.cos(angle), math.sin(angle)
        transform = RigidTransform((c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0), self.centre_mm)
        rows, columns = self.shape
        du, dv = self.spacing_mm
        origin = (
            self.detector_offset_mm[0] - 0.5 * (columns - 1) * du,
            self.source_detector_distance_mm - self.source_object_distance_mm,
            self.detector_offset_mm[1] - 0.5 * (rows - 1) * dv,
        )
        return DetectorGeometry(
            transform.point((0.0, -self.source_object_distance_mm, 0.0)),
            transform.point(origin),
            transform.direction((1.0, 0.0, 0.0)),
            transform.direction((0.0, 0.0, 1.0)),
            self.spacing_mm,
            self.shape,
        )

    def batch(self, angles_rad) -> BatchedGeometry:
        return BatchedGeometry(tuple(self.view(float(a)) for a in angles_rad))

    def ray_derivatives(self, angle_rad: float) -> tuple[tuple[Vector3, Vector3], ...]:
        """(d source, d endpoint) of every pixel of the view at ``angle_rad``
        for SOD, SDD, column offset and row offset, in world mm per mm
        (geometry.py:688-701): the source sits at -SOD along the rotated y
        axis, the detector plane at SDD - SOD, the offsets move the panel
        along the rotated x and world z axes."""
        angle = finite_tuple((float(angle_rad),), "view angle", minimum=None)[0]
        x = (math.cos(angle), math.sin(angle), 0.0)
        y = (-math.sin(angle), math.cos(angle), 0.0)
        minus_y = (-y[0], -y[1], -y[2])
        zero = (0.0, 0.0, 0.0)
        return ((minus_y, minus_y), (zero, y), (zero, x), (zero, (0.0, 0.0, 1.0)))


def _skew(w) -> Matrix3:
    x, y, z = w
    return (0.0, -z, y, z, 0.0, -x, -y, x, 0.0)


def _se3_coefficients(theta2: float) -> tuple[float, float, float]:
    """sin(t)/t, (1-cos t)/t^2, (t-sin t)/t^3 with series near zero."""
    if not math.isfinite(theta2):
        raise ContractError("rotation norm exceeds the finite geometry range")
    if theta2 < 1e-2:
        a = 1 - theta2 / 6 + theta2**2 / 120 - theta2**3 / 5040
        b = 0.5 - theta2 / 24 + theta2**2 / 720 - theta2**3 / 40320
        c = 1 / 6 - theta2 / 120 + theta2**2 / 5040 - theta2**3 / 362880
        return a, b, c
    t = math.sqrt(theta2)
    return math.sin(t) / t, (1 - math.cos(t)) / theta2, (t - math.sin(t)) / (theta2 * t)


def compose_pose(anchor: RigidTransform, increment) -> RigidTransform:
    """``anchor @ exp(xi^)``; xi = (tx,ty,tz,rx,ry,rz) in mm / radians."""
    xi = finite_tuple(tuple(float(x) for x in increment), "pose increment", minimum=None, length=6)
    rho, w = xi[:3], xi[3:]
    omega = _skew(w)
    omega2 = matmul(omega, omega)
    a, b, c = _se3_coefficients(dot(w, w))
    rotation = tuple(IDENTITY[i] + a * omega[i] + b * omega2[i] for i in range(9))
    velocity = tuple(IDENTITY[i] + b * omega[i] + c * omega2[i] for i in range(9))
    return anchor.compose(RigidTransform(cast(Matrix3, rotation), matvec(velocity, rho)))
