"""Test scenes as plain numbers, shared by oracle and product tests.

Each scene restates a fixture of the reference's own tests (cited) so the
oracle and the HIP path are compared on the same inputs.
"""

from __future__ import annotations

import math

import numpy as np

from oracle import oracle as O

IDENTITY12 = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0)


def compose(anchor12, xi):
    """anchor @ exp(xi^) with closed-form SE(3) exponential (FP64, host)."""
    rho = np.asarray(xi[:3], dtype=np.float64)
    w = np.asarray(xi[3:], dtype=np.float64)
    theta2 = float(w @ w)
    hat = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]], dtype=np.float64)
    if theta2 < 1e-2:
        a = 1 - theta2 / 6 + theta2**2 / 120 - theta2**3 / 5040
        b = 0.5 - theta2 / 24 + theta2**2 / 720 - theta2**3 / 40320
        c = 1 / 6 - theta2 / 120 + theta2**2 / 5040 - theta2**3 / 362880
    else:
        t = math.sqrt(theta2)
        a, b, c = math.sin(t) / t, (1 - math.cos(t)) / theta2, (t - math.sin(t)) / (theta2 * t)
    rot = np.eye(3) + a * hat + b * hat @ hat
    vel = np.eye(3) + b * hat + c * hat @ hat
    ra = np.asarray(anchor12[:9], dtype=np.float64).reshape(3, 3)
    ta = np.asarray(anchor12[9:], dtype=np.float64)
    return tuple((ra @ rot).reshape(-1)) + tuple(ra @ (vel @ rho) + ta)


def oblique_scene():
    """tests/python/warp_cpu/test_projection_fp32.py:24-68 (5x6x7 oriented grid)."""
    angle = 0.19
    grid = O.Grid(
        (5, 6, 7),
        (1.2, 0.8, 1.1),
        (-3.2, -2.1, -1.7),
        (math.cos(angle), -math.sin(angle), 0.0, math.sin(angle), math.cos(angle), 0.0, 0.0, 0.0, 1.0),
    )
    sources = np.array(
        [(-12.0, 0.17, 0.37), (12.0, 0.23, 0.48), (-0.7, -12.0, 0.29), (0.31, 0.41, -12.0), (-12.0, 30.0, 30.0)]
    )
    endpoints = np.array(
        [(12.0, 0.81, 0.63), (-12.0, 0.89, 0.91), (0.2, 12.0, 0.86), (0.72, 0.83, 12.0), (12.0, 30.0, 30.0)]
    )
    pose = compose(IDENTITY12, (0.07, -0.03, 0.02, 0.013, -0.021, 0.009))
    values = np.asarray(
        [
            0.06 + 0.013 * x + 0.019 * y + 0.011 * z + 0.003 * x * y + 0.002 * y * z
            for z in range(5)
            for y in range(6)
            for x in range(7)
        ],
        dtype=np.float32,
    )
    return grid, sources, endpoints, pose, values


def workflow_scene(views: int = 2):
    """tests/python/warp_cpu/test_projector_workflow.py:27-42 (3x3x4 grid, 2x3 panels)."""
    grid = O.Grid((3, 3, 4), (0.7, 0.8, 0.9), (-1.0, -1.0, -1.0))
    base = O.Panel((-8.0, 0.1, -0.2), (8.0, -0.3, -0.4), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.2, 0.2), (2, 3))
    if views == 2:
        panels = [base, O.Panel(base.source, (8.0, -0.24, -0.31), base.u, base.v, base.spacing, base.shape)]
    else:
        panels = [
            O.Panel(base.source, (8.0, -0.3 + i * 0.03, -0.4), base.u, base.v, base.spacing, base.shape)
            for i in range(views)
        ]
    field = np.linspace(0.1, 0.3, 36, dtype=np.float32)
    return grid, panels, field


def line_panel(source: float = -10.0, detector: float = 10.0):
    """tests/python/test_projection_reference.py:16-24 (single pixel on the x axis)."""
    return O.Panel((source, 0.0, 0.0), (detector, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (1.0, 1.0), (1, 1))


def cone_panels(grid: O.Grid, views: int, rows: int, cols: int, sod=1000.0, sdd=1500.0, pitch=None, offset=0.0):
    """Circular cone views about world z centred on the grid (geometry.py:617-682)."""
    nx, ny, nz = grid.xyz
    centre = [grid.origin[i] + ((grid.xyz[i] - 1) / 2) * grid.spacing[i] for i in range(3)]
    if pitch is None:
        diag = math.hypot(nx * grid.spacing[0], ny * grid.spacing[1], nz * grid.spacing[2])
        pitch = 1.1 * diag * sdd / (sod - diag / 2) / max(rows, cols)
    panels = []
    for k in range(views):
        a = offset + 2 * math.pi * k / views
        c, s = math.cos(a), math.sin(a)

        def rot(p):
            return (c * p[0] - s * p[1] + centre[0], s * p[0] + c * p[1] + centre[1], p[2] + centre[2])

        origin = (-0.5 * (cols - 1) * pitch, sdd - sod, -0.5 * (rows - 1) * pitch)
        panels.append(
            O.Panel(rot((0.0, -sod, 0.0)), rot(origin), (c, s, 0.0), (0.0, 0.0, 1.0), (pitch, pitch), (rows, cols))
        )
    return panels


def smooth_field(grid: O.Grid, seed: int = 0) -> np.ndarray:
    """Positive smooth field plus a seeded perturbation (float32)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = grid.xyz
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = 0.02 + 0.01 * np.sin(0.3 * x + 0.1) * np.cos(0.2 * y) + 0.005 * np.cos(0.25 * z)
    return (base + 0.002 * rng.random(base.shape)).astype(np.float32)


GATHER_CASES = ("nonsquare", "bundle", "clamped")


def gather_scene(case: str):
    """tests/python/warp_cpu/test_projection_adjoint_gather.py:26-69: a 2x5
    panel at a 0.1 rad yaw, a six-ray bundle with axis-parallel, oblique and
    missing rays, and a clamped 1x2x3 grid.  Returns (grid, geometry kwargs,
    pose)."""
    if case == "clamped":
        grid = O.Grid((1, 2, 3), (1.0, 0.7, 1.2), (-1.0, 0.0, 0.0))
        sources = np.array([(-4.0, -0.35, 0.0), (-4.0, 0.9, 0.0), (0.1, 0.1, 0.0), (-4.0, 9.0, 0.0)])
        endpoints = np.array([(4.0, -0.35, 0.0), (4.0, 0.9, 0.0), (0.7, 0.3, 0.0), (4.0, 9.0, 0.0)])
        return grid, dict(rays=(sources, endpoints)), IDENTITY12
    grid = O.Grid((2, 3, 4), (1.2, 0.8, 1.1), (-1.8, -0.8, -0.55))
    pose = compose(IDENTITY12, (0.07, -0.03, 0.02, 0.01, -0.02, 0.005))
    if case == "bundle":
        sources = np.array(
            [(-8.0, 0.1, 0.2), (8.0, 0.1, 0.2), (-8.0, 9.0, 9.0), (0.1, 0.2, 0.3), (-1.0, -5.0, 0.2), (-0.7, 0.4, -8.0)]
        )
        endpoints = np.array(
            [(8.0, 0.8, 0.7), (-8.0, 0.8, 0.7), (8.0, 9.0, 9.0), (0.3, 0.6, 0.7), (-1.0, 5.0, 0.3), (0.2, 0.4, 8.0)]
        )
        return grid, dict(rays=(sources, endpoints)), pose
    panel = O.Panel(
        (-8.0, 0.17, 0.23), (8.0, -0.43, -0.37), (math.sin(0.1), math.cos(0.1), 0.0), (0.0, 0.0, 1.0), (0.19, 0.23), (2, 5)
    )
    return grid, dict(panels=[panel]), pose


def mixed_panels(grid: O.Grid, rows: int = 9, cols: int = 11):
    """A cone panel, a parallel panel and a cylindrical panel of one shape,
    looking at the grid from three directions (geometry.py:306-431)."""
    nx, ny, nz = grid.xyz
    centre = [grid.origin[i] + ((grid.xyz[i] - 1) / 2) * grid.spacing[i] for i in range(3)]
    extent = max(nx * grid.spacing[0], ny * grid.spacing[1], nz * grid.spacing[2])
    pitch = 1.3 * extent / max(rows, cols)
    cone = O.Panel((centre[0], centre[1] - 60.0, centre[2]),
                   (centre[0] - 0.5 * (cols - 1) * pitch, centre[1] + 40.0, centre[2] - 0.5 * (rows - 1) * pitch),
                   (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), (pitch, pitch), (rows, cols))
    a = 0.35
    direction = (math.cos(a), math.sin(a), 0.0)  # unit beam direction
    u = (-math.sin(a), math.cos(a), 0.0)
    origin = tuple(centre[i] + 30.0 * direction[i] - 0.5 * (cols - 1) * 0.8 * pitch * u[i]
                   - (0.5 * (rows - 1) * 0.8 * pitch if i == 2 else 0.0) for i in range(3))
    parallel = dict(origin_mm=origin, direction=direction, length_mm=70.0, u=u, v=(0.0, 0.0, 1.0),
                    spacing_mm=(0.8 * pitch, 0.8 * pitch), shape=(rows, cols))
    # cylinder of radius 90 about the vertical axis through the source
    source = (centre[0] - 50.0, centre[1] + 10.0, centre[2])
    b = -0.25  # pixel (0,0) sits at this angle from the +x direction around the axis
    radius = 90.0
    origin_c = (source[0] + radius * math.cos(b), source[1] + radius * math.sin(b), centre[2] - 0.5 * (rows - 1) * pitch)
    u_c = (-math.sin(b), math.cos(b), 0.0)
    curved = dict(source_mm=source, origin_mm=origin_c, u=u_c, v=(0.0, 0.0, 1.0),
                  spacing_mm=(pitch * 1.4, pitch), shape=(rows, cols))
    return cone, parallel, curved


def parallel_segment(p: dict, row: int, col: int):
    """geometry.py:349-354 restated: endpoint = origin + col du u + row dv v,
    source = endpoint - length * direction."""
    du, dv = p["spacing_mm"]
    end = tuple(p["origin_mm"][i] + col * du * p["u"][i] + row * dv * p["v"][i] for i in range(3))
    return tuple(end[i] - p["length_mm"] * p["direction"][i] for i in range(3)), end


def curved_segment(p: dict, row: int, col: int):
    """geometry.py:408-431 restated: the pixel rotates about the axis through
    the source along v by the arc angle col * du / radius."""
    offset = [p["origin_mm"][i] - p["source_mm"][i] for i in range(3)]
    along = sum(offset[i] * p["v"][i] for i in range(3))
    radial = [offset[i] - along * p["v"][i] for i in range(3)]
    radius = math.sqrt(sum(x * x for x in radial))
    angle = col * p["spacing_mm"][0] / radius
    end = tuple(p["origin_mm"][i] + (math.cos(angle) - 1.0) * radial[i] + math.sin(angle) * radius * p["u"][i]
                + row * p["spacing_mm"][1] * p["v"][i] for i in range(3))
    return tuple(p["source_mm"]), end


# tests/python/warp_cpu/test_projection_shared_face.py:91-153: rays that walk
# the cells of small grids along and across the axes, seeds scaled by 1 and
# 1000 (cell sums that carry across shared faces), overflow-scale seeds
# accumulated onto an opposing destination, and subnormal-scale seeds whose
# every increment the GPU rule flushes to zero.
SHARED_FACE_DIRECTIONS = ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, -1))


def shared_face_scene(direction):
    """A 5x5x5 unit grid and three rays through its centre along +-direction
    (plus one that misses), field linspace(0.01, 0.1)."""
    centre = np.asarray([2.0, 2.0, 2.0])
    displacement = np.asarray(direction, dtype=np.float64) * 10.0
    source, target = centre - displacement, centre + displacement
    sources = np.array([source, target, source + 100.0])
    endpoints = np.array([target, source, target + 100.0])
    grid = O.Grid((5, 5, 5), (1.0, 1.0, 1.0))
    return grid, sources, endpoints, np.linspace(0.01, 0.1, grid.voxels, dtype=np.float32)


def overflow_scene(reverse: bool):
    """Grid (2, 2, 3) at (2, 1, 1) mm and one ray along x: each grid sample
    it passes collects about 2 mm x seed."""
    source, target = np.array([[-3.0, 0.0, 0.0]]), np.array([[7.0, 0.0, 0.0]])
    if reverse:
        source, target = target, source
    grid = O.Grid((2, 2, 3), (2.0, 1.0, 1.0))
    return grid, source, target, np.linspace(0.01, 0.1, grid.voxels, dtype=np.float32)


def subnormal_scene():
    """Grid (2, 2, 3) of unit spacing and one ray along x at y = z = 0.5: with
    a seed of 1e-35 every increment factor x weight is below 2^-126."""
    grid = O.Grid((2, 2, 3), (1.0, 1.0, 1.0))
    return (grid, np.array([[-2.0, 0.5, 0.5]]), np.array([[5.0, 0.5, 0.5]]),
            np.linspace(0.01, 0.1, grid.voxels, dtype=np.float32))
