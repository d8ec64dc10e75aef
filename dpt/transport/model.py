"""Transport records: material cells, detectors, statuses and the transport spec.

Host-side mirror of the reference's transport model
(python/dpt/transport/model.py:28-580): same record names, field meanings,
units (mm, keV, mm^-1) and validation messages for the analogue Woodcock path.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from enum import IntEnum
from typing import Literal

from ..contracts import ContractError
from ..geometry import DetectorGeometry, RigidTransform, cross, dot
from ..volumes import GridSpec

ELECTRON_REST_ENERGY_KEV = 510.99895069


class TransportError(ContractError):
    """A transport physical, storage or execution contract was violated (model.py:28)."""


class IncompleteHistoryError(RuntimeError):
    """A resource guard stopped a history; the entire estimate is invalid (model.py:32)."""


class HistoryStatus(IntEnum):
    """Each launched history writes exactly one terminal status (model.py:36-45)."""

    ESCAPED = 0
    ABSORBED = 1
    EVENT_BUDGET = 2
    CROSSING_BUDGET = 3
    NUMERICAL_FAILURE = 4
    ENERGY_SUPPORT = 5
    ANGLE_BUDGET = 6


NUMERICAL_MASK = 1 << HistoryStatus.NUMERICAL_FAILURE
ENERGY_MASK = 1 << HistoryStatus.ENERGY_SUPPORT
BUDGET_MASK = (
    (1 << HistoryStatus.EVENT_BUDGET) | (1 << HistoryStatus.CROSSING_BUDGET) | (1 << HistoryStatus.ANGLE_BUDGET)
)


def _physical(value, name: str, *, positive: bool = False, nonnegative: bool = False) -> float:
    if isinstance(value, bool) or not isinstance(value, (int, float)):
        raise TransportError(f"{name} must be a real number")
    value = float(value)
    if not math.isfinite(value):
        raise TransportError(f"{name} must be finite")
    if positive and not value > 0:
        raise TransportError(f"{name} must be positive")
    if nonnegative and value < 0:
        raise TransportError(f"{name} must be nonnegative")
    return value


def _positive_integer(value, name: str) -> int:
    if type(value) is not int or value <= 0:
        raise TransportError(f"{name} must be a positive integer")
    return value


@dataclass(frozen=True)
class MaterialGrid:
    """Axis-aligned cell faces; shape is (nz, ny, nx), labels x fastest (model.py:72-150)."""

    origin_mm: tuple[float, float, float]
    spacing_mm: tuple[float, float, float]
    shape: tuple[int, int, int]
    materials: int

    def __post_init__(self) -> None:
        for name in ("origin_mm", "spacing_mm", "shape"):
            object.__setattr__(self, name, tuple(getattr(self, name)))
        if len(self.origin_mm) != 3 or len(self.spacing_mm) != 3 or len(self.shape) != 3:
            raise TransportError("origin, spacing and shape must each have three components")
        _positive_integer(self.materials, "materials")
        for count, step, origin in zip(self.shape_xyz, self.spacing_mm, self.origin_mm):
            _positive_integer(count, "grid extent")
            _physical(step, "cell spacing", positive=True)
            _physical(origin, "cell origin")
            end = origin + count * step
            if not math.isfinite(end) or origin + step == origin or end - step == end:
                raise TransportError("cell faces must remain distinguishable in binary64")
        if math.prod(self.shape) >= 2**31:
            raise TransportError("the flattened material grid exceeds signed 32-bit indexing")

    @classmethod
    def from_grid_spec(cls, grid: GridSpec, *, materials: int) -> MaterialGrid:
        """Constant material cells matching an axis-aligned sample grid's support
        (the first face lies half a spacing below the first sample centre)."""
        origin = tuple(c - s / 2 for c, s in zip(grid.origin_mm, grid.spacing_mm))
        return cls(origin, tuple(grid.spacing_mm), tuple(grid.shape), materials)

    @property
    def shape_xyz(self) -> tuple[int, int, int]:
        return self.shape[2], self.shape[1], self.shape[0]

    @property
    def upper_mm(self) -> tuple[float, float, float]:
        return tuple(o + n * s for o, n, s in zip(self.origin_mm, self.shape_xyz, self.spacing_mm))


@dataclass(frozen=True)
class PlanarDetector:
    """Outward-facing z plane with lower pixel faces (model.py:166-281)."""

    lower_xy_mm: tuple[float, float]
    spacing_xy_mm: tuple[float, float]
    shape: tuple[int, int]  # width, height
    z_mm: float

    def __post_init__(self) -> None:
        for name in ("lower_xy_mm", "spacing_xy_mm", "shape"):
            object.__setattr__(self, name, tuple(getattr(self, name)))
        _physical(self.z_mm, "detector z")
        for count, step, lower in zip(self.shape, self.spacing_xy_mm, self.lower_xy_mm):
            _positive_integer(count, "detector extent")
            _physical(step, "pixel spacing", positive=True)
            _physical(lower, "detector origin")
        if self.pixels >= 2**31:
            raise TransportError("the flattened detector exceeds signed 32-bit indexing")

    @classmethod
    def from_detector_geometry(cls, detector: DetectorGeometry) -> PlanarDetector:
        if tuple(detector.u) != (1.0, 0.0, 0.0) or tuple(detector.v) != (0.0, 1.0, 0.0):
            raise TransportError("PlanarDetector conversion requires positive x/y detector axes")
        return cls(
            (detector.origin_mm[0] - detector.spacing_mm[0] / 2, detector.origin_mm[1] - detector.spacing_mm[1] / 2),
            detector.spacing_mm,
            (detector.shape[1], detector.shape[0]),
            detector.origin_mm[2],
        )

    @property
    def pixels(self) -> int:
        return self.shape[0] * self.shape[1]

    @property
    def image_shape(self) -> tuple[int, int]:
        return self.shape[1], self.shape[0]

    @property
    def lower_corner_mm(self) -> tuple[float, float, float]:
        return (self.lower_xy_mm[0], self.lower_xy_mm[1], self.z_mm)

    @property
    def u(self) -> tuple[float, float, float]:
        return (1.0, 0.0, 0.0)

    @property
    def v(self) -> tuple[float, float, float]:
        return (0.0, 1.0, 0.0)

    @property
    def normal(self) -> tuple[float, float, float]:
        return (0.0, 0.0, 1.0)

    def source_distance(self, position) -> float:
        return position[2] - self.z_mm


@dataclass(frozen=True)
class OrientedDetector:
    """An existing detector geometry with half-open pixel footprints (model.py:284-349)."""

    geometry: DetectorGeometry

    def __post_init__(self) -> None:
        if not isinstance(self.geometry, DetectorGeometry):
            raise TransportError("OrientedDetector.geometry must be a DetectorGeometry")

    @property
    def lower_corner_mm(self) -> tuple[float, float, float]:
        g = self.geometry
        return tuple(g.origin_mm[i] - 0.5 * (g.spacing_mm[0] * g.u[i] + g.spacing_mm[1] * g.v[i]) for i in range(3))

    @property
    def u(self):
        return self.geometry.u

    @property
    def v(self):
        return self.geometry.v

    @property
    def normal(self) -> tuple[float, float, float]:
        n = cross(self.u, self.v)
        offset = tuple(self.geometry.source_mm[i] - self.geometry.origin_mm[i] for i in range(3))
        return tuple(-x for x in n) if dot(offset, n) > 0 else tuple(n)

    def source_distance(self, position) -> float:
        corner = self.lower_corner_mm
        return dot(tuple(position[i] - corner[i] for i in range(3)), self.normal)

    @property
    def spacing_xy_mm(self) -> tuple[float, float]:
        return self.geometry.spacing_mm

    @property
    def shape(self) -> tuple[int, int]:
        return self.geometry.shape[1], self.geometry.shape[0]

    @property
    def pixels(self) -> int:
        return self.geometry.pixels

    @property
    def image_shape(self) -> tuple[int, int]:
        return tuple(self.geometry.shape)


@dataclass(frozen=True)
class TransportSpec:
    """Physical scope and launch-resource policy (model.py:420-580).

    Supported here: the analogue estimator with isotropic-elastic or
    free-electron Compton scattering, photon-count or keV scoring, the
    legacy and packed Klein-Nishina samplers, and an object pose.
    """

    grid: MaterialGrid
    detector: PlanarDetector | OrientedDetector
    energy_kev: float
    coefficient_provenance: str
    coefficient_energies_kev: tuple[float, ...] = ()
    scattering_law: Literal["isotropic-elastic", "free-electron-compton"] = "isotropic-elastic"
    scoring: Literal["photon-count", "energy-kev"] = "photon-count"
    max_angle_trials: int = 4096
    max_events: int = 4096
    max_crossings: int = 65536
    block_dim: int = 256
    estimator: Literal["analogue"] = "analogue"
    energy_support_policy: Literal["runtime-guard"] = "runtime-guard"
    object_pose: RigidTransform = field(default_factory=RigidTransform)
    compton_sampler: Literal["legacy-rejection", "packed-uniform-rejection"] = "legacy-rejection"

    def __post_init__(self) -> None:
        object.__setattr__(self, "coefficient_energies_kev", tuple(float(e) for e in self.coefficient_energies_kev))
        if not isinstance(self.grid, MaterialGrid):
            raise TransportError("grid must be a MaterialGrid")
        if not isinstance(self.detector, (PlanarDetector, OrientedDetector)):
            raise TransportError("detector must be a PlanarDetector or OrientedDetector")
        if not isinstance(self.object_pose, RigidTransform):
            raise TransportError("object_pose must be a RigidTransform")
        _physical(self.energy_kev, "source energy", positive=True)
        if type(self.coefficient_provenance) is not str or not self.coefficient_provenance.strip():
            raise TransportError("the supplied coefficients require a provenance identifier")
        for x in (self.grid.origin_mm[0], self.grid.upper_mm[0]):
            for y in (self.grid.origin_mm[1], self.grid.upper_mm[1]):
                for z in (self.grid.origin_mm[2], self.grid.upper_mm[2]):
                    distance = self.detector.source_distance(self.object_pose.point((x, y, z)))
                    if not math.isfinite(distance) or distance >= 0:
                        raise TransportError(
                            "the detector plane must lie above/on the outgoing side of the entire material grid"
                        )
        nodes = self.energy_nodes
        if any(b <= a for a, b in zip(nodes, nodes[1:])):
            raise TransportError("coefficient energies must be strictly increasing")
        if not nodes[0] <= self.energy_kev <= nodes[-1]:
            raise TransportError("coefficient energies must contain the source energy")
        if self.scattering_law not in ("isotropic-elastic", "free-electron-compton"):
            raise TransportError("unsupported scattering law")
        if self.scattering_law != "isotropic-elastic" and len(nodes) < 2:
            raise TransportError("Compton scattering needs energy-dependent coefficient tables")
        if self.scoring not in ("photon-count", "energy-kev"):
            raise TransportError("unsupported detector score")
        if self.estimator != "analogue":
            raise TransportError("unsupported transport estimator")
        if self.energy_support_policy != "runtime-guard":
            raise TransportError("unsupported coefficient energy support policy")
        if self.compton_sampler not in ("legacy-rejection", "packed-uniform-rejection"):
            raise TransportError("unsupported Compton sampler execution policy")
        for name in ("max_angle_trials", "max_events", "max_crossings"):
            _positive_integer(getattr(self, name), name)
            if getattr(self, name) >= 2**31:
                raise TransportError(f"{name} must fit signed 32-bit integers")
        if self.block_dim not in (64, 128, 256):
            raise TransportError("block_dim must be 64, 128 or 256")

    @property
    def energy_nodes(self) -> tuple[float, ...]:
        return self.coefficient_energies_kev or (float(self.energy_kev),)

    def with_execution(self, **changes) -> TransportSpec:
        return replace(self, **changes)
