"""CPU oracle for the cone-beam projector -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The product
package ``dpt`` never imports it.

Two layers:

* ``cpu_ref`` (``oracle/cpu_ref.c``, loaded through ctypes): the operation-for-
  operation restatement of the reference FP32 path
  (``kernels/projection_fp32.py:334-617``) and FP64 path
  (``kernels/projection.py:802-1237``), multithreaded over rays.
* Pure NumPy/Python restatements that do not share code with ``cpu_ref``:
  - ``explicit_fp32_matrix``: the independent rounded-coordinate interpolation
    matrix of ``tests/python/warp_cpu/test_projection_fp32.py:76-119``;
  - ``integrate_segment``: exact piecewise two-point Gauss integral (or an
    independent midpoint rule) of the trilinear field,
    ``validation/projection.py:57-251``;
  - ``sample_reference``: nested linear interpolation,
    ``validation/projection.py:79-114``;
  - ``ellipsoid_chord``: analytic ellipsoid optical depth, ``phantoms.py:56-97``.

Geometry is passed as plain numbers (grid shape/spacing/origin/orientation,
per-view source and detector vectors) so the oracle never depends on the
product's geometry classes.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "libcpu_ref.so"
_LIB = None

IDENTITY_POSE = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0)
RAY_COORDINATES = 7  # the C oracle's per-ray source/endpoint derivative mode


def build() -> Path:
    """Compile ``cpu_ref.c`` with the committed Makefile (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


class _Plan(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("views", ctypes.c_int32),
        ("rows", ctypes.c_int32),
        ("cols", ctypes.c_int32),
        ("samples", ctypes.c_int32),
        ("bundle", ctypes.c_int32),
        ("g", ctypes.c_double * 9),
        ("grid_origin", ctypes.c_double * 3),
        ("panels", ctypes.c_void_p),
        ("ray_sources", ctypes.c_void_p),
        ("ray_endpoints", ctypes.c_void_p),
    ]


def _lib():
    global _LIB
    if _LIB is None:
        if not _LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(_LIB_PATH))
        p = ctypes.c_void_p
        for name in ("ref_forward_f32", "ref_forward_f64"):
            getattr(lib, name).argtypes = [ctypes.POINTER(_Plan), p, p, p, p]
            getattr(lib, name).restype = None
        for name in ("ref_vjp_f32", "ref_vjp_f64"):
            getattr(lib, name).argtypes = [ctypes.POINTER(_Plan), p, p, p, p, p, p, ctypes.c_int, p]
            getattr(lib, name).restype = None
        lib.ref_forward_gauss_f64.argtypes = [ctypes.POINTER(_Plan), p, p, p, p]
        lib.ref_forward_gauss_f64.restype = None
        lib.ref_vjp_gauss_f64.argtypes = [ctypes.POINTER(_Plan), p, p, p, p, p, p, ctypes.c_int, p]
        lib.ref_vjp_gauss_f64.restype = None
        lib.ref_gather_f64.argtypes = [ctypes.POINTER(_Plan), p, p, p, ctypes.c_int, p]
        lib.ref_gather_f64.restype = None
        lib.ref_threads.restype = ctypes.c_int
        lib.ref_set_threads.argtypes = [ctypes.c_int]
        lib.ref_set_threads.restype = None
        _LIB = lib
    return _LIB


def threads() -> int:
    return int(_lib().ref_threads())


def set_threads(n: int) -> None:
    """Thread count for later oracle calls in this process."""
    _lib().ref_set_threads(int(n))


@dataclass(frozen=True)
class Grid:
    """Grid description: shape (nz,ny,nx), spacing (x,y,z), origin = first centre."""

    shape: tuple[int, int, int]
    spacing: tuple[float, float, float]
    origin: tuple[float, float, float] = (0.0, 0.0, 0.0)
    orientation: tuple[float, ...] = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)

    @property
    def xyz(self) -> tuple[int, int, int]:
        return (self.shape[2], self.shape[1], self.shape[0])

    @property
    def voxels(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]

    def object_to_grid(self) -> list[float]:
        # projection.py:573-576: G[i][j] = orientation[3j+i] / spacing[i]
        return [self.orientation[3 * j + i] / self.spacing[i] for i in range(3) for j in range(3)]


@dataclass(frozen=True)
class Panel:
    """Flat cone panel: pixel (r,c) centre = origin + c*du*u + r*dv*v."""

    source: tuple[float, float, float]
    origin: tuple[float, float, float]
    u: tuple[float, float, float]
    v: tuple[float, float, float]
    spacing: tuple[float, float]  # (column, row)
    shape: tuple[int, int]  # (rows, columns)

    def packed(self) -> list[float]:
        du, dv = self.spacing
        return [
            *self.source,
            *self.origin,
            *(x * du for x in self.u),
            *(x * dv for x in self.v),
        ]

    def pixel(self, row: int, column: int) -> tuple[float, float, float]:
        du, dv = self.spacing
        return tuple(
            self.origin[i] + column * du * self.u[i] + row * dv * self.v[i] for i in range(3)
        )


class Scene:
    """A grid plus either panels (one per view) or explicit rays ((rays, 3) for
    one view, (views, rays, 3) for one bundle per view)."""

    def __init__(
        self,
        grid: Grid,
        samples: int,
        panels: list[Panel] | None = None,
        rays: tuple[np.ndarray, np.ndarray] | None = None,
    ) -> None:
        self.grid, self.samples = grid, int(samples)
        plan = _Plan()
        plan.nx, plan.ny, plan.nz = grid.xyz
        plan.samples = self.samples
        for i, value in enumerate(grid.object_to_grid()):
            plan.g[i] = value
        for i in range(3):
            plan.grid_origin[i] = grid.origin[i]
        if panels is not None:
            shapes = {p.shape for p in panels}
            if len(shapes) != 1:
                raise ValueError("panels must share (rows, columns)")
            self._panels = np.ascontiguousarray([p.packed() for p in panels], dtype=np.float64)
            plan.views, (plan.rows, plan.cols) = len(panels), panels[0].shape
            plan.bundle = 0
            plan.panels = self._panels.ctypes.data
        else:
            sources, endpoints = rays
            sources = np.asarray(sources, dtype=np.float64)
            views = sources.shape[0] if sources.ndim == 3 else 1  # (views, rays, 3): one bundle per view
            self._sources = np.ascontiguousarray(sources, dtype=np.float64).reshape(-1, 3)
            self._endpoints = np.ascontiguousarray(endpoints, dtype=np.float64).reshape(-1, 3)
            plan.views, plan.rows, plan.cols = views, 1, self._sources.shape[0] // views
            plan.bundle = 1
            plan.ray_sources = self._sources.ctypes.data
            plan.ray_endpoints = self._endpoints.ctypes.data
        self.plan = plan
        self.views, self.rows, self.cols = plan.views, plan.rows, plan.cols

    @property
    def image_shape(self) -> tuple[int, int, int]:
        return (self.views, self.rows, self.cols)

    def _poses(self, poses) -> np.ndarray:
        poses = np.asarray(IDENTITY_POSE if poses is None else poses, dtype=np.float64)
        if poses.ndim == 1:
            poses = np.broadcast_to(poses, (self.views, 12))
        return np.ascontiguousarray(poses, dtype=np.float64)

    def forward(self, field, poses=None, *, precision: str = "fp32", integration: str = "midpoint"):
        """Return (image (V,R,C), status (V,) int32).  ``integration="cell_gauss"``
        (FP64 only) is the exact cell-Gauss rule of kernels/projection.py:522-746."""
        field = np.ascontiguousarray(field, dtype=np.float32).reshape(-1)
        if field.size != self.grid.voxels:
            raise ValueError("field size does not match grid")
        poses = self._poses(poses)
        dtype = np.float32 if precision == "fp32" else np.float64
        out = np.zeros(self.image_shape, dtype=dtype)
        status = np.zeros(self.views, dtype=np.int32)
        fn = _lib().ref_forward_f32 if precision == "fp32" else _lib().ref_forward_f64
        if integration == "cell_gauss":
            if precision != "fp64":
                raise ValueError("cell-Gauss integration is FP64 only")
            fn = _lib().ref_forward_gauss_f64
        fn(ctypes.byref(self.plan), field.ctypes.data, poses.ctypes.data, out.ctypes.data,
           status.ctypes.data)
        return out, status

    def vjp(
        self,
        field,
        cotangent,
        poses=None,
        *,
        precision: str = "fp32",
        volume: bool = True,
        pose: bool = True,
        components: int = 12,
        per_ray: bool = False,
        integration: str = "midpoint",
    ):
        """Return (grad_field FP64 (voxels,) | None, grad_pose (V,C) | None,
        per_ray (V*R*C, C) | None, status)."""
        field = np.ascontiguousarray(field, dtype=np.float32).reshape(-1)
        poses = self._poses(poses)
        seed_dtype = np.float32 if precision == "fp32" else np.float64
        seeds = np.ascontiguousarray(cotangent, dtype=seed_dtype).reshape(-1)
        if seeds.size != self.views * self.rows * self.cols:
            raise ValueError("cotangent size does not match image")
        grad_field = np.zeros(self.grid.voxels, dtype=np.float64) if volume else None
        grad_pose = np.zeros((self.views, components), dtype=np.float64) if pose else None
        rays = np.zeros((seeds.size, components), dtype=np.float64) if per_ray else None
        status = np.zeros(self.views, dtype=np.int32)
        fn = _lib().ref_vjp_f32 if precision == "fp32" else _lib().ref_vjp_f64
        if integration == "cell_gauss":
            if precision != "fp64":
                raise ValueError("cell-Gauss integration is FP64 only")
            fn = _lib().ref_vjp_gauss_f64

        def ptr(a):
            return None if a is None else a.ctypes.data

        fn(ctypes.byref(self.plan), field.ctypes.data, poses.ctypes.data, seeds.ctypes.data,
           ptr(grad_field), ptr(grad_pose), ptr(rays), components, status.ctypes.data)
        return grad_field, grad_pose, rays, status

    def ray_vjp(self, field, cotangent, poses=None, *, precision: str = "fp32", integration: str = "midpoint"):
        """Per ray: cotangent x d(optical depth)/d(world source xyz, world
        endpoint xyz) (kernels/projection_fp32.py:570-583,
        kernels/projection.py:1184-1200).  Returns ((V*R*C, 6), status)."""
        field = np.ascontiguousarray(field, dtype=np.float32).reshape(-1)
        poses = self._poses(poses)
        seed_dtype = np.float32 if precision == "fp32" else np.float64
        seeds = np.ascontiguousarray(cotangent, dtype=seed_dtype).reshape(-1)
        rays = np.zeros((seeds.size, 6), dtype=np.float64)
        status = np.zeros(self.views, dtype=np.int32)
        fn = _lib().ref_vjp_f32 if precision == "fp32" else _lib().ref_vjp_f64
        if integration == "cell_gauss":
            fn = _lib().ref_vjp_gauss_f64
        fn(ctypes.byref(self.plan), field.ctypes.data, poses.ctypes.data, seeds.ctypes.data, None, None,
           rays.ctypes.data, RAY_COORDINATES, status.ctypes.data)
        return rays, status

    def gather(self, cotangent, poses=None, *, previous=None, precision: str = "float64"):
        """Deterministic FP64 adjoint gather (kernels/projection_adjoint.py:73-170).

        Returns (out (voxels,) in ``precision``, status (V,) int32).  With
        ``previous`` the call accumulates: out = dtype(total + float64(previous)).
        """
        poses = self._poses(poses)
        seeds = np.ascontiguousarray(cotangent, dtype=np.float64).reshape(-1)
        if seeds.size != self.views * self.rows * self.cols:
            raise ValueError("cotangent size does not match image")
        out = np.zeros(self.grid.voxels, dtype=np.float64)
        if previous is not None:
            out[:] = np.asarray(previous).reshape(-1).astype(np.float64)
        status = np.zeros(self.views, dtype=np.int32)
        _lib().ref_gather_f64(ctypes.byref(self.plan), poses.ctypes.data, seeds.ctypes.data, out.ctypes.data,
                              int(previous is not None), status.ctypes.data)
        if precision == "float32":
            out = out.astype(np.float32)
            if not np.isfinite(out).all():
                status[0] |= 2
        return out, status

    def rays_grid(self, poses=None):
        """Per-ray grid-frame (origin, direction, length) in the reference's
        FP64 operation order (kernels/projection.py:79-102), for the
        voxel-driven restatement below."""
        poses = self._poses(poses)
        g = self.grid.object_to_grid()
        out = []
        for view in range(self.views):
            pose = [float(x) for x in poses[view]]
            for pixel in range(self.rows * self.cols):
                if self.plan.bundle:
                    index = view * self.rows * self.cols + pixel
                    src = [float(x) for x in self._sources[index]]
                    end = [float(x) for x in self._endpoints[index]]
                else:
                    pn = [float(x) for x in self._panels[view]]
                    row, col = divmod(pixel, self.cols)
                    src = pn[0:3]
                    end = [pn[3 + i] + float(col) * pn[6 + i] + float(row) * pn[9 + i] for i in range(3)]
                wsr = [src[i] - pose[9 + i] for i in range(3)]
                wd = [end[i] - src[i] for i in range(3)]
                osrc = [pose[i] * wsr[0] + pose[3 + i] * wsr[1] + pose[6 + i] * wsr[2] for i in range(3)]
                odir = [pose[i] * wd[0] + pose[3 + i] * wd[1] + pose[6 + i] * wd[2] for i in range(3)]
                rel = [osrc[i] - self.grid.origin[i] for i in range(3)]
                origin = [g[3 * i] * rel[0] + g[3 * i + 1] * rel[1] + g[3 * i + 2] * rel[2] for i in range(3)]
                direction = [g[3 * i] * odir[0] + g[3 * i + 1] * odir[1] + g[3 * i + 2] * odir[2]
                             for i in range(3)]
                length = math.sqrt(wd[0] * wd[0] + wd[1] * wd[1] + wd[2] * wd[2])
                out.append((origin, direction, length))
        return out


def gather_voxel_driven(scene: Scene, cotangent, poses=None) -> np.ndarray:
    """Voxel-driven restatement of get_gather (kernels/projection_adjoint.py:
    79-166, midpoint, finite inputs): every voxel visits every ray, sample and
    corner in order and keeps the addends whose clamped corner index is the
    voxel (point_contribution, :33-44).  Pure Python; tiny grids only."""
    nx, ny, nz = scene.grid.xyz
    shape = (nx, ny, nz)
    seeds = [float(x) for x in np.asarray(cotangent, dtype=np.float64).reshape(-1)]
    rays = scene.rays_grid(poses)
    samples = scene.samples

    def interval(o, d):
        lower, upper = 0.0, 1.0
        for a in range(3):
            low, high = -0.5, float(shape[a]) - 0.5
            if d[a] == 0.0:
                if o[a] < low or o[a] > high:
                    upper = -1.0
            else:
                first, last = (low - o[a]) / d[a], (high - o[a]) / d[a]
                if first > last:
                    first, last = last, first
                lower = max(lower, first) if first > lower else lower
                upper = last if last < upper else upper
        return lower, upper

    out = np.zeros(nx * ny * nz, dtype=np.float64)
    for voxel in range(nx * ny * nz):
        total = 0.0
        for ray, (o, d, length) in enumerate(rays):
            seed = seeds[ray]
            if seed == 0.0:
                continue
            lower, upper = interval(o, d)
            span = upper - lower
            if span <= 0.0:
                continue
            step = span / float(samples)
            volume_factor = seed * length * span / float(samples)
            for k in range(samples):
                pos = lower + (float(k) + 0.5) * step
                p = [o[a] + pos * d[a] for a in range(3)]
                if not all(-0.5 <= p[a] <= float(shape[a]) - 0.5 for a in range(3)):
                    continue
                low, frac = [], []
                for a in range(3):
                    c = min(max(p[a], 0.0), float(shape[a] - 1))
                    lo = min(int(math.floor(c)), max(shape[a] - 2, 0))
                    low.append(lo)
                    frac.append(c - float(lo))
                for corner in range(8):
                    idx = [min(low[a] + ((corner >> a) & 1), shape[a] - 1) for a in range(3)]
                    if (idx[2] * ny + idx[1]) * nx + idx[0] != voxel:
                        continue
                    contribution = volume_factor
                    for a in range(3):
                        contribution *= frac[a] if (corner >> a) & 1 else 1.0 - frac[a]
                    total += contribution
        out[voxel] = total
    return out


# --------------------------------------------------------------------------
# Independent restatements (pure Python / NumPy, no shared code with cpu_ref)
# --------------------------------------------------------------------------


def pose_coordinates(grid: Grid, pose, point) -> tuple[float, float, float]:
    """World point -> grid index coordinates (validation/projection.py:21-37)."""
    pose = IDENTITY_POSE if pose is None else pose
    displacement = [point[i] - pose[9 + i] for i in range(3)]
    obj = [sum(pose[3 * j + i] * displacement[j] for j in range(3)) for i in range(3)]
    return tuple(
        sum(grid.orientation[3 * j + i] * (obj[j] - grid.origin[j]) for j in range(3))
        / grid.spacing[i]
        for i in range(3)
    )


def clip_segment(grid: Grid, pose, source, endpoint):
    """validation/projection.py:57-76: (start, direction, lower, upper, length)."""
    start = pose_coordinates(grid, pose, source)
    end = pose_coordinates(grid, pose, endpoint)
    direction = tuple(end[i] - start[i] for i in range(3))
    lower, upper = 0.0, 1.0
    for axis, extent in enumerate(grid.xyz):
        if direction[axis] == 0.0:
            if start[axis] < -0.5 or start[axis] > extent - 0.5:
                return start, direction, 0.0, 0.0, 0.0
        else:
            faces = [
                (-0.5 - start[axis]) / direction[axis],
                (extent - 0.5 - start[axis]) / direction[axis],
            ]
            lower = max(lower, min(faces))
            upper = min(upper, max(faces))
    length = math.sqrt(sum((endpoint[i] - source[i]) ** 2 for i in range(3)))
    return start, direction, lower, upper, length


def sample_reference(values, grid: Grid, point) -> float:
    """Nested linear interpolation with half-cell clamp (validation/projection.py:79-114)."""
    xyz = grid.xyz
    if any(point[i] < -0.5 or point[i] > xyz[i] - 0.5 for i in range(3)):
        return 0.0
    q = [min(max(point[i], 0.0), xyz[i] - 1.0) for i in range(3)]
    lo = [math.floor(x) for x in q]
    hi = [min(lo[i] + 1, xyz[i] - 1) for i in range(3)]
    fraction = [q[i] - lo[i] for i in range(3)]
    planes = []
    for z in (lo[2], hi[2]):
        rows = []
        for y in (lo[1], hi[1]):
            a = values[(z * xyz[1] + y) * xyz[0] + lo[0]]
            b = values[(z * xyz[1] + y) * xyz[0] + hi[0]]
            rows.append(a + (b - a) * fraction[0])
        planes.append(rows[0] + (rows[1] - rows[0]) * fraction[1])
    return planes[0] + (planes[1] - planes[0]) * fraction[2]


def integrate_segment(values, grid: Grid, pose, source, endpoint, midpoint_samples=None) -> float:
    """Exact piecewise Gauss (default) or midpoint integral of one segment.

    Restates validation/projection.py:204-251.
    """
    values = [float(x) for x in np.asarray(values, dtype=np.float64).reshape(-1)]
    start, direction, lower, upper, length = clip_segment(grid, pose, source, endpoint)
    if upper <= lower:
        return 0.0

    def sample(t):
        return sample_reference(values, grid, tuple(start[i] + t * direction[i] for i in range(3)))

    if midpoint_samples is not None:
        step = (upper - lower) / midpoint_samples
        return length * step * math.fsum(sample(lower + (i + 0.5) * step) for i in range(midpoint_samples))
    breaks = {lower, upper}
    for axis, extent in enumerate(grid.xyz):
        if direction[axis] != 0:
            for centre in range(extent):
                t = (centre - start[axis]) / direction[axis]
                if lower < t < upper:
                    breaks.add(t)
    knots = sorted(breaks)
    total = []
    for lo, hi in zip(knots[:-1], knots[1:]):
        c, r = (lo + hi) / 2, (hi - lo) / 2
        off = r / math.sqrt(3)
        total.append(r * (sample(c - off) + sample(c + off)))
    return length * math.fsum(total)


def _fp32_ray_samples(grid: Grid, pose, source, endpoint, samples: int):
    """One ray of the rounded-coordinate FP32 walk: (physical step, [(voxel,
    weight) per inside sample and corner]) or None for a miss.  FP32 ray
    coordinates and clipping, FP32 midpoints via an exact-product FMA, FP32
    weights (tests/python/warp_cpu/test_projection_fp32.py:71-119)."""
    f32 = np.float32

    def fma32(a, b, c):
        return f32(float(a) * float(b) + float(c))

    shape = grid.xyz
    origin, direction, _, _, length = clip_segment(grid, pose, source, endpoint)
    origin = np.asarray(origin, dtype=f32)
    direction = np.asarray(direction, dtype=f32)
    lower, upper = f32(0), f32(1)
    for axis in range(3):
        low, high = f32(-0.5), f32(shape[axis] - 0.5)
        velocity = direction[axis]
        if velocity == 0:
            if origin[axis] < low or origin[axis] > high:
                return None
        else:
            first = f32(f32(low - origin[axis]) / velocity)
            last = f32(f32(high - origin[axis]) / velocity)
            first, last = min(first, last), max(first, last)
            lower, upper = max(lower, first), min(upper, last)
    if upper <= lower:
        return None
    step = f32(f32(upper - lower) / f32(samples))
    physical = f32(f32(length) * step)
    terms = []
    for index in range(samples):
        position = fma32(f32(index) + f32(0.5), step, lower)
        point = [fma32(position, direction[a], origin[a]) for a in range(3)]
        if any(point[a] < -0.5 or point[a] > shape[a] - 0.5 for a in range(3)):
            continue
        vals = [f32(min(max(point[a], 0), shape[a] - 1)) for a in range(3)]
        cell = [min(math.floor(vals[a]), max(shape[a] - 2, 0)) for a in range(3)]
        fractions = [f32(vals[a] - f32(cell[a])) for a in range(3)]
        for corner in range(8):
            idx = [min(cell[a] + ((corner >> a) & 1), shape[a] - 1) for a in range(3)]
            factors = [fractions[a] if (corner >> a) & 1 else f32(1) - fractions[a] for a in range(3)]
            weight = f32(f32(factors[0] * factors[1]) * factors[2])
            terms.append(((idx[2] * shape[1] + idx[1]) * shape[0] + idx[0], weight))
    return physical, terms


def explicit_fp32_matrix(grid: Grid, segments, pose, samples: int) -> np.ndarray:
    """Rounded-coordinate FP32 interpolation matrix, (rays, voxels) in FP64.

    Restates tests/python/warp_cpu/test_projection_fp32.py:71-119: each entry
    is RN32(physical_step * weight) summed in FP64.
    """
    f32 = np.float32
    matrix = np.zeros((len(segments), grid.voxels), dtype=np.float64)
    for ray, (source, endpoint) in enumerate(segments):
        walked = _fp32_ray_samples(grid, pose, source, endpoint, samples)
        if walked is None:
            continue
        physical, terms = walked
        for voxel, weight in terms:
            matrix[ray, voxel] += float(f32(physical * weight))
    return matrix


def explicit_fp32_scatter(grid: Grid, segments, pose, samples: int, seeds, *, flush: bool = True) -> np.ndarray:
    """Per-sample FP32 scatter of cotangents, summed in FP64 (voxels,).

    Restates the per-sample reference scatter of
    tests/python/warp_cpu/test_projection_cell_aggregation.py:22-48 with the
    GPU increment rule of kernels/projection_fp32.py:240-252: factor =
    RN32(seed * physical_step), increment RN32(factor * weight), flushed to a
    signed zero below 2^-126 when ``flush`` (the reference's GPU path;
    ``flush=False`` is its FP64-destination scatter, :212-235).
    """
    f32 = np.float32
    tiny = float(np.finfo(np.float32).tiny)
    out = np.zeros(grid.voxels, dtype=np.float64)
    for (source, endpoint), seed in zip(segments, seeds):
        seed = f32(seed)
        if seed == 0:
            continue
        walked = _fp32_ray_samples(grid, pose, source, endpoint, samples)
        if walked is None:
            continue
        physical, terms = walked
        with np.errstate(over="ignore"):
            factor = f32(seed * physical)
            for voxel, weight in terms:
                increment = f32(factor * weight)
                if flush and abs(float(increment)) < tiny:
                    increment = f32(0.0)
                out[voxel] += float(increment)
    return out


def ellipsoid_chord(axes, mu, centre, source, target) -> float:
    """Analytic optical depth of an axis-aligned ellipsoid (phantoms.py:56-97)."""
    d = [target[i] - source[i] for i in range(3)]
    o = [(source[i] - centre[i]) / axes[i] for i in range(3)]
    s = [d[i] / axes[i] for i in range(3)]
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
    lower = max(-h, along)
    upper = min(h, n + along)
    return mu * max(0.0, upper - lower) * (math.hypot(*d) / n)


def relative_l2(actual, expected) -> float:
    a = np.asarray(actual, dtype=np.float64).ravel()
    e = np.asarray(expected, dtype=np.float64).ravel()
    denom = float(np.linalg.norm(e))
    diff = float(np.linalg.norm(a - e))
    return diff / denom if denom > 0 else diff

