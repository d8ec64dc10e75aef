"""Analytic ellipsoids, seeded ellipsoid phantoms and circular acquisitions.

``Ellipsoid`` / ``sample_ellipsoids`` restate python/dpt/phantoms.py:25-155
(analytic chord; additive sampling at voxel centres).  ``chest_phantom`` is the
seeded synthetic volume of the benchmark workloads: a body ellipsoid
(mu = 0.020/mm), two lung ellipsoids (0.005/mm) and five bone ellipsoids
(0.050/mm), painted in that order with *later ellipsoid overrides* semantics
(the reference's additive sampler cannot express lungs inside a body with
non-negative coefficients), plus Gaussian noise (sigma 1e-4) and a clamp at
zero so the physical-field check (finite, >= 0) holds.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .contracts import ContractError, finite_scalar
from .geometry import IDENTITY, BatchedGeometry, ConeBeamIntrinsics, Matrix3, Vector3, matvec, transpose, vector3
from .volumes import GridSpec

BODY_MU = 0.020
LUNG_MU = 0.005
BONE_MU = 0.050
NOISE_SIGMA = 1e-4


@dataclass(frozen=True)
class Ellipsoid:
    """mu(x) = coefficient where sum((R^T (x-centre) / axes)^2) <= 1."""

    axes_mm: Vector3
    attenuation_mm_inverse: float
    centre_mm: Vector3 = (0.0, 0.0, 0.0)
    orientation: Matrix3 = IDENTITY

    def __post_init__(self) -> None:
        object.__setattr__(self, "axes_mm", vector3(self.axes_mm, "axes_mm"))
        if min(self.axes_mm) <= 0:
            raise ContractError("ellipsoid axes must be positive")
        object.__setattr__(self, "centre_mm", vector3(self.centre_mm, "centre_mm"))
        object.__setattr__(
            self,
            "attenuation_mm_inverse",
            finite_scalar(self.attenuation_mm_inverse, "attenuation_mm_inverse", minimum=0),
        )

    def optical_depth(self, source_mm, target_mm) -> float:
        """coefficient * |segment intersect ellipsoid| (phantoms.py:56-97)."""
        d = tuple(target_mm[i] - source_mm[i] for i in range(3))
        local = matvec(transpose(self.orientation), tuple(source_mm[i] - self.centre_mm[i] for i in range(3)))
        direction = matvec(transpose(self.orientation), d)
        o = [local[i] / self.axes_mm[i] for i in range(3)]
        s = [direction[i] / self.axes_mm[i] for i in range(3)]
        n = math.hypot(*s)
        if n == 0:
            return 0.0
        u = [x / n for x in s]
        along = math.fsum(o[i] * u[i] for i in range(3))
        nearest = [o[i] - along * u[i] for i in range(3)]
        r2 = 1.0 - math.fsum(x * x for x in nearest)
        if r2 <= 0:
            return 0.0
        h = math.sqrt(r2)
        lower, upper = max(-h, along), min(h, n + along)
        return self.attenuation_mm_inverse * max(0.0, upper - lower) * (math.hypot(*d) / n)

    def mask(self, grid: GridSpec, z: int) -> np.ndarray:
        """Boolean (ny, nx) membership of one z slab of voxel centres."""
        nz, ny, nx = grid.shape
        x = np.arange(nx, dtype=np.float64)[None, :] * grid.spacing_mm[0]
        y = np.arange(ny, dtype=np.float64)[:, None] * grid.spacing_mm[1]
        zz = z * grid.spacing_mm[2]
        o = grid.orientation
        slab = [grid.origin_mm[a] + o[3 * a] * x + o[3 * a + 1] * y + o[3 * a + 2] * zz for a in range(3)]
        inv = transpose(self.orientation)
        distance = np.zeros((ny, nx), dtype=np.float64)
        for a in range(3):
            c = sum(inv[3 * a + k] * (slab[k] - self.centre_mm[k]) for k in range(3)) / self.axes_mm[a]
            distance = distance + c * c
        return distance <= 1.0


def sample_ellipsoids(grid: GridSpec, ellipsoids, *, dtype: str = "float32") -> np.ndarray:
    """values[z,y,x] = sum_j mu_j at voxel centres (additive; phantoms.py:100-155)."""
    if dtype not in ("float32", "float64"):
        raise ContractError("phantom dtype must be float32 or float64")
    models = tuple(ellipsoids)
    result = np.zeros(grid.shape, dtype=dtype)
    for z in range(grid.shape[0]):
        for model in models:
            result[z] += model.mask(grid, z) * model.attenuation_mm_inverse
    return result


def centred_grid(shape_zyx, spacing_xyz=(1.0, 1.0, 1.0)) -> GridSpec:
    """Grid whose support is centred on the world origin."""
    nz, ny, nx = shape_zyx
    h = tuple(float(s) for s in spacing_xyz)
    origin = (-(nx - 1) * h[0] / 2, -(ny - 1) * h[1] / 2, -(nz - 1) * h[2] / 2)
    return GridSpec((int(nz), int(ny), int(nx)), h, origin)


def chest_ellipsoids(grid: GridSpec, seed: int) -> list[Ellipsoid]:
    """Seeded body, 2 lungs and 5 bones scaled to the grid's physical extent."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = grid.xyz
    half = (nx * grid.spacing_mm[0] / 2, ny * grid.spacing_mm[1] / 2, nz * grid.spacing_mm[2] / 2)
    c0 = grid.centre_mm()

    def jitter(scale):
        return 1.0 + scale * (2.0 * rng.random() - 1.0)

    def place(fx, fy, fz):
        return (c0[0] + fx * half[0], c0[1] + fy * half[1], c0[2] + fz * half[2])

    def axes(ax, ay, az):
        return (ax * half[0] * jitter(0.08), ay * half[1] * jitter(0.08), az * half[2] * jitter(0.08))

    body = Ellipsoid(axes(0.86, 0.62, 0.90), BODY_MU, place(0.02 * jitter(1), 0.02 * jitter(1), 0.0))
    lungs = [
        Ellipsoid(axes(0.28, 0.34, 0.62), LUNG_MU, place(side * 0.36 * jitter(0.1), -0.05 * jitter(0.2), 0.05 * jitter(1)))
        for side in (-1.0, 1.0)
    ]
    bone_layout = [
        ((0.0, 0.46, 0.0), (0.09, 0.09, 0.80)),   # spine
        ((0.0, -0.52, 0.05), (0.06, 0.04, 0.42)), # sternum
        ((-0.62, 0.10, 0.0), (0.05, 0.05, 0.55)), # left rib band
        ((0.62, 0.10, 0.0), (0.05, 0.05, 0.55)),  # right rib band
        ((0.0, 0.30, -0.70), (0.36, 0.12, 0.08)), # pelvis/girdle slab
    ]
    bones = [
        Ellipsoid(
            axes(*a),
            BONE_MU,
            place(c[0] + 0.03 * (2 * rng.random() - 1), c[1] + 0.03 * (2 * rng.random() - 1), c[2]),
        )
        for c, a in bone_layout
    ]
    return [body, *lungs, *bones]


def chest_phantom(grid: GridSpec, seed: int = 0, *, noise_sigma: float = NOISE_SIGMA) -> np.ndarray:
    """Seeded float32 (nz,ny,nx) attenuation in mm^-1 (overrides, noise, clamp >= 0)."""
    models = chest_ellipsoids(grid, seed)
    rng = np.random.default_rng(seed + 0x5EED)
    out = np.zeros(grid.shape, dtype=np.float32)
    for z in range(grid.shape[0]):
        slab = np.zeros(grid.shape[1:], dtype=np.float64)
        for model in models:
            slab = np.where(model.mask(grid, z), model.attenuation_mm_inverse, slab)
        if noise_sigma > 0:
            slab = slab + rng.normal(0.0, noise_sigma, size=slab.shape)
        out[z] = np.maximum(slab, 0.0).astype(np.float32)
    return out


def chest_labels(values: np.ndarray) -> np.ndarray:
    """Three-material labels: 0 air (mu <= 0.001), 1 soft (<= 0.035), 2 bone."""
    labels = np.full(values.shape, 2, dtype=np.uint8)
    labels[values <= 0.035] = 1
    labels[values <= 0.001] = 0
    return labels


def coverage_pitch(
    grid: GridSpec,
    sod: float,
    sdd: float,
    angles,
    shape: tuple[int, int],
    *,
    margin: float = 1.05,
) -> float:
    """Square pitch whose panel covers every support corner in every view.

    Rule of benchmarks/fullscan_inputs.py:139-176: project the eight support
    corners perspective-correctly onto each panel, take the largest absolute
    column/row offset from the panel centre, pitch = margin * 2 * max / N.
    """
    corners = np.asarray(grid.support_corners_mm())
    probe = ConeBeamIntrinsics(sod, sdd, shape, (1.0, 1.0), centre_mm=grid.centre_mm())
    worst = 0.0
    for angle in angles:
        panel = probe.view(float(angle))
        source = np.asarray(panel.source_mm)
        u, v = np.asarray(panel.u), np.asarray(panel.v)
        centre = np.asarray(panel.origin_mm) + (
            (shape[1] - 1) * panel.spacing_mm[0] * u + (shape[0] - 1) * panel.spacing_mm[1] * v
        ) / 2
        axis = centre - source
        distance = np.linalg.norm(axis)
        normal = axis / distance
        depth = (corners - source) @ normal
        if np.any(depth <= 0):
            raise ContractError("volume support reaches the source plane")
        projected = source + (corners - source) * (distance / depth)[:, None]
        offsets = np.column_stack(((projected - centre) @ u, (projected - centre) @ v))
        worst = max(worst, float(np.max(np.abs(offsets))))
    return margin * 2.0 * worst / max(shape)


def circular_geometry(
    grid: GridSpec,
    views: int,
    shape: tuple[int, int],
    *,
    sod: float = 1000.0,
    sdd: float = 1500.0,
    pitch: float | None = None,
) -> tuple[BatchedGeometry, float]:
    """Full circular orbit, angles 2*pi*k/views, centred on the grid."""
    angles = [2.0 * math.pi * k / views for k in range(views)]
    if pitch is None:
        pitch = coverage_pitch(grid, sod, sdd, angles, shape)
    intrinsics = ConeBeamIntrinsics(sod, sdd, shape, (pitch, pitch), centre_mm=grid.centre_mm())
    return intrinsics.batch(angles), pitch


def standard_angles() -> dict[str, float]:
    """AP, lateral and oblique source angles (python/dpt/imaging.py:170)."""
    return {"ap": math.pi, "lateral": math.pi / 2, "oblique": 3 * math.pi / 4}
