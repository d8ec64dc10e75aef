"""Flat cone, parallel and cylindrical panels in one batch on the GPU.

The plan expands each view's panel record into world segments on the device
(kernels/batched_geometry.py:9-39); the oracle receives segments generated
independently in Python from the reference's formulas
(python/dpt/geometry.py:306-431, restated in tests/scenes.py).  The bars are
those of the flat-panel parity tests: FP32 images 1e-6 relative L2, FP64
1e-12, gradients 1e-5 (FP32); FP64 pose gradients 1e-9 and volume gradients
1e-7 (an FP64 sum rounded once to FP32).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import scenes

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt.geometry import BatchedGeometry, CurvedDetectorGeometry, DetectorGeometry, ParallelGeometry  # noqa: E402
from dpt.torch import Projector  # noqa: E402
from dpt.volumes import GridSpec  # noqa: E402


def _setup(samples=96):
    grid = O.Grid((12, 13, 14), (1.0, 0.9, 1.1), (-6.5, -5.4, -7.15))
    cone, parallel, curved = scenes.mixed_panels(grid)
    geometry = BatchedGeometry((
        DetectorGeometry(cone.source, cone.origin, cone.u, cone.v, cone.spacing, cone.shape),
        ParallelGeometry(**parallel),
        CurvedDetectorGeometry(**curved),
    ))
    rows, cols = cone.shape
    sources, endpoints = [], []
    for view in range(3):
        src, end = [], []
        for r in range(rows):
            for c in range(cols):
                if view == 0:
                    s, e = cone.source, cone.pixel(r, c)
                elif view == 1:
                    s, e = scenes.parallel_segment(parallel, r, c)
                else:
                    s, e = scenes.curved_segment(curved, r, c)
                src.append(s)
                end.append(e)
        sources.append(src)
        endpoints.append(end)
    ref = O.Scene(grid, samples, rays=(np.asarray(sources), np.asarray(endpoints)))
    spec = GridSpec(grid.shape, grid.spacing, grid.origin, grid.orientation)
    return grid, spec, geometry, ref, samples


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_mixed_panels_match_the_oracle(precision):
    print(f"case mixed panels {precision}", flush=True)
    grid, spec, geometry, ref, samples = _setup()
    values = scenes.smooth_field(grid, 7)
    module = Projector(grid=spec, geometry=geometry, samples_per_ray=samples, precision=precision)
    packed = np.asarray([scenes.compose(scenes.IDENTITY12, (0.4, -0.3, 0.2, 0.03, -0.02, 0.025))] * 3)
    field = torch.tensor(values, device="cuda", requires_grad=True)
    pose = torch.tensor(packed, device="cuda", requires_grad=True)  # (views, 12) packed poses
    image = module(field, packed_pose=pose)
    assert tuple(image.shape) == (3, 9, 11)
    arith = "fp32" if precision == "fp32" else "fp64"
    expected, status = ref.forward(values, packed, precision=arith)
    assert status.tolist() == [0, 0, 0]
    got = image.detach().cpu().numpy()
    assert np.count_nonzero(expected) > 0.5 * expected.size
    assert _rel(got, expected) <= (1e-6 if precision == "fp32" else 1e-12)
    dtype = torch.float32 if precision == "fp32" else torch.float64
    seeds = np.random.default_rng(3).uniform(-1.0, 1.0, expected.size)
    seed = torch.tensor(seeds.reshape(tuple(image.shape)), dtype=dtype, device="cuda")
    g_field, g_pose = torch.autograd.grad(image, (field, pose), seed)
    g_ref, p_ref, _, _ = ref.vjp(values, seeds, packed, precision=arith)
    # the FP64 path returns its FP64 volume sum rounded once to FP32: 1e-7
    assert _rel(g_field.cpu().numpy(), g_ref) <= (1e-5 if precision == "fp32" else 1e-7)
    assert _rel(g_pose.cpu().numpy(), p_ref) <= (1e-5 if precision == "fp32" else 1e-9)


def test_parallel_panel_of_a_constant_field_measures_chords():
    """A constant field along axis-parallel rays: every ray that crosses the
    whole grid integrates value x the grid's support length along x."""
    grid = GridSpec((6, 7, 8), (1.0, 1.0, 1.0), (-3.5, -3.0, -2.5))
    panel = ParallelGeometry((20.0, -1.5, -1.0), (1.0, 0.0, 0.0), 40.0, (0.0, 1.0, 0.0), (0.0, 0.0, 1.0),
                             (0.5, 0.5), (4, 5))
    module = Projector(grid=grid, geometry=panel, samples_per_ray=64, precision="fp64")
    image = module(torch.full(grid.shape, 0.25, device="cuda"))
    torch.testing.assert_close(image, torch.full((1, 4, 5), 0.25 * 8.0, dtype=torch.float64, device="cuda"),
                               rtol=1e-13, atol=0)


def test_gather_on_mixed_panels_is_bitwise_the_oracle():
    """The deterministic gather over device-generated segments of cone,
    parallel and cylindrical panels (projected pixel windows for the cone and
    parallel views, every pixel for the cylindrical one) repeats bit for bit
    and equals the oracle's gather over the independently generated segments
    to 1e-12 (the cylindrical endpoints use the device's cos/sin, which may
    differ from the host's in the last bit)."""
    grid, spec, geometry, ref, samples = _setup(samples=31)
    module = Projector(grid=spec, geometry=geometry, samples_per_ray=samples, precision="fp64")
    seeds = np.random.default_rng(4).uniform(-1.0, 1.0, 3 * 9 * 11)
    packed = np.asarray([scenes.compose(scenes.IDENTITY12, (0.2, 0.1, -0.3, 0.01, 0.02, -0.01))] * 3)
    out = module.adjoint_gather(torch.tensor(seeds.reshape(3, 9, 11), device="cuda"),
                                packed_pose=torch.tensor(packed, device="cuda"))
    expected, status = ref.gather(seeds, packed)
    assert status.tolist() == [0, 0, 0] and np.count_nonzero(expected) > 0
    again = module.adjoint_gather(torch.tensor(seeds.reshape(3, 9, 11), device="cuda"),
                                  packed_pose=torch.tensor(packed, device="cuda"))
    assert torch.equal(out, again)
    np.testing.assert_allclose(out.cpu().numpy().reshape(-1), expected, rtol=1e-12,
                               atol=1e-14 * np.abs(expected).max())


def _parallel_views(grid: O.Grid, rows: int, cols: int):
    """Parallel panels at several beam directions: along the x axis, against
    the y axis (beam components of exactly zero), and three oblique ones, one
    with a tilted row direction; each panel covers part of the grid."""
    centre = [grid.origin[i] + ((grid.xyz[i] - 1) / 2) * grid.spacing[i] for i in range(3)]
    out = []
    for k, (a, tilt) in enumerate([(0.0, 0.0), (-math.pi / 2, 0.0), (0.35, 0.0), (2.2, 0.3), (-1.1, -0.45)]):
        direction = (math.cos(a) * math.cos(tilt), math.sin(a) * math.cos(tilt), math.sin(tilt))
        u = (-math.sin(a), math.cos(a), 0.0)
        v = (-math.cos(a) * math.sin(tilt), -math.sin(a) * math.sin(tilt), math.cos(tilt))
        pitch = (0.55 + 0.1 * k, 0.6)
        shift = (0.3 * k - 0.6, 0.2 * k - 0.4)
        origin = tuple(centre[i] + 25.0 * direction[i] - (0.5 * (cols - 1) * pitch[0] + shift[0]) * u[i]
                       - (0.5 * (rows - 1) * pitch[1] + shift[1]) * v[i] for i in range(3))
        out.append(ParallelGeometry(origin, direction, 50.0, u, v, pitch, (rows, cols)))
    return out


def test_gather_on_parallel_panels_is_bitwise_the_oracle():
    """Parallel panels made from panel records take the gather's affine
    pixel windows (every other voxel skips most pixels); the sums stay those
    of the full loop bit for bit: FP64 and FP32 destinations equal the
    oracle's serial scatter over host-generated segments (geometry.py:349-354
    of python/dpt restated), with an off-identity pose."""
    print("case gather parallel panels", flush=True)
    grid = O.Grid((16, 15, 17), (0.9, 1.0, 0.8), (-7.0, -7.2, -6.1))
    rows, cols = 13, 17
    panels = _parallel_views(grid, rows, cols)
    sources, endpoints = [], []
    for p in panels:
        segs = [scenes.parallel_segment(dict(origin_mm=p.origin_mm, direction=p.direction, length_mm=p.length_mm,
                                             u=p.u, v=p.v, spacing_mm=p.spacing_mm), r, c)
                for r in range(rows) for c in range(cols)]
        sources.append([s for s, _ in segs])
        endpoints.append([e for _, e in segs])
    samples = 37
    ref = O.Scene(grid, samples, rays=(np.asarray(sources), np.asarray(endpoints)))
    spec = GridSpec(grid.shape, grid.spacing, grid.origin, grid.orientation)
    module = Projector(grid=spec, geometry=BatchedGeometry(tuple(panels)), samples_per_ray=samples,
                       precision="fp64")
    views = len(panels)
    seeds = np.random.default_rng(6).uniform(-1.0, 1.0, views * rows * cols)
    packed = np.asarray([scenes.compose(scenes.IDENTITY12, (0.1, -0.2, 0.15, 0.02, -0.01, 0.03))] * views)
    cot = torch.tensor(seeds.reshape(views, rows, cols), device="cuda")
    pose = torch.tensor(packed, device="cuda")
    out = module.adjoint_gather(cot, packed_pose=pose)
    expected, status = ref.gather(seeds, packed)
    assert status.tolist() == [0] * views
    assert np.count_nonzero(expected) > 0.5 * expected.size
    assert out.cpu().numpy().reshape(-1).tobytes() == expected.tobytes()
    single = module.adjoint_gather(cot, packed_pose=pose, volume_gradient_precision="float32")
    assert single.cpu().numpy().reshape(-1).tobytes() == expected.astype(np.float32).tobytes()


def test_gather_on_cone_panel_records_matches_the_oracle():
    """Flat cone panels inside a mixed batch reach the gather as panel
    records and take the projective pixel windows; two cone views at
    different angles beside a parallel one, off-identity pose: equal to the
    oracle's serial scatter over host-generated segments to 1e-12 per voxel
    (a dropped window would miss whole rays, far above that), bit-identical
    on repeat."""
    print("case gather cone panel records", flush=True)
    grid = O.Grid((14, 16, 15), (1.0, 0.9, 1.1), (-6.5, -6.75, -7.7))
    cone, parallel, _ = scenes.mixed_panels(grid, rows=12, cols=13)
    rows, cols = cone.shape
    centre = [grid.origin[i] + ((grid.xyz[i] - 1) / 2) * grid.spacing[i] for i in range(3)]
    a = 0.9  # second cone panel: the first rotated about the vertical axis through the centre
    def turn(p):
        x, y = p[0] - centre[0], p[1] - centre[1]
        return (centre[0] + math.cos(a) * x - math.sin(a) * y, centre[1] + math.sin(a) * x + math.cos(a) * y, p[2])
    u2 = (math.cos(a) * cone.u[0] - math.sin(a) * cone.u[1], math.sin(a) * cone.u[0] + math.cos(a) * cone.u[1], cone.u[2])
    cone2 = O.Panel(turn(cone.source), turn(cone.origin), u2, cone.v, cone.spacing, cone.shape)
    panels = (DetectorGeometry(cone.source, cone.origin, cone.u, cone.v, cone.spacing, cone.shape),
              DetectorGeometry(cone2.source, cone2.origin, cone2.u, cone2.v, cone2.spacing, cone2.shape),
              ParallelGeometry(**parallel))
    sources, endpoints = [], []
    for view in range(3):
        src, end = [], []
        for r in range(rows):
            for c in range(cols):
                if view < 2:
                    p = (cone, cone2)[view]
                    s_, e_ = p.source, p.pixel(r, c)
                else:
                    s_, e_ = scenes.parallel_segment(parallel, r, c)
                src.append(s_)
                end.append(e_)
        sources.append(src)
        endpoints.append(end)
    samples = 29
    ref = O.Scene(grid, samples, rays=(np.asarray(sources), np.asarray(endpoints)))
    spec = GridSpec(grid.shape, grid.spacing, grid.origin, grid.orientation)
    module = Projector(grid=spec, geometry=BatchedGeometry(panels), samples_per_ray=samples, precision="fp64")
    seeds = np.random.default_rng(8).uniform(-1.0, 1.0, 3 * rows * cols)
    packed = np.asarray([scenes.compose(scenes.IDENTITY12, (-0.2, 0.1, 0.25, 0.015, 0.02, -0.02))] * 3)
    cot = torch.tensor(seeds.reshape(3, rows, cols), device="cuda")
    pose = torch.tensor(packed, device="cuda")
    out = module.adjoint_gather(cot, packed_pose=pose)
    assert torch.equal(out, module.adjoint_gather(cot, packed_pose=pose))
    expected, status = ref.gather(seeds, packed)
    assert status.tolist() == [0, 0, 0] and np.count_nonzero(expected) > 0.5 * expected.size
    np.testing.assert_allclose(out.cpu().numpy().reshape(-1), expected, rtol=1e-12,
                               atol=1e-14 * np.abs(expected).max())


def test_gather_on_panel_records_at_size_matches_adjoint_identity_and_scatter():
    """96^3 grid, 12 parallel views (beam angles over half a turn, some
    tilted) and 4 flat cone views in one batch of 80 x 88 pixels, 192
    samples: the culled gather satisfies <A x, c> = <x, A^T c> against the
    forward to 1e-10 and matches the atomic volume VJP to 1e-5, without the
    oracle."""
    print("case gather panel records 96^3 x 16 x 80x88 S=192", flush=True)
    n, rows, cols = 96, 80, 88
    grid = O.Grid((n, n, n), (1.0, 1.0, 1.0), (-47.5, -47.5, -47.5))
    views = []
    for k in range(12):
        a, tilt = math.pi * k / 12, (0.2 if k % 3 == 1 else 0.0)
        direction = (math.cos(a) * math.cos(tilt), math.sin(a) * math.cos(tilt), math.sin(tilt))
        u = (-math.sin(a), math.cos(a), 0.0)
        v = (-math.cos(a) * math.sin(tilt), -math.sin(a) * math.sin(tilt), math.cos(tilt))
        pitch = 1.25
        origin = tuple(60.0 * direction[i] - 0.5 * (cols - 1) * pitch * u[i] - 0.5 * (rows - 1) * pitch * v[i]
                       for i in range(3))
        views.append(ParallelGeometry(origin, direction, 120.0, u, v, (pitch, pitch), (rows, cols)))
    for k in range(4):
        a = 0.4 + math.pi * k / 2
        d = (math.cos(a), math.sin(a), 0.0)
        u = (-math.sin(a), math.cos(a), 0.0)
        pitch = 1.9
        source = tuple(-250.0 * d[i] for i in range(3))
        origin = tuple(150.0 * d[i] - 0.5 * (cols - 1) * pitch * u[i] - (0.5 * (rows - 1) * pitch if i == 2 else 0.0)
                       for i in range(3))
        views.append(DetectorGeometry(source, origin, u, (0.0, 0.0, 1.0), (pitch, pitch), (rows, cols)))
    spec = GridSpec(grid.shape, grid.spacing, grid.origin, grid.orientation)
    module = Projector(grid=spec, geometry=BatchedGeometry(tuple(views)), samples_per_ray=192, precision="fp64")
    packed = torch.tensor(np.asarray([scenes.compose(scenes.IDENTITY12, (0.6, -0.4, 0.2, 0.02, -0.01, 0.015))] * 16),
                          device="cuda")
    rng = np.random.default_rng(31)
    seeds = torch.tensor(rng.uniform(-1.0, 1.0, (16, rows, cols)), device="cuda")
    values = torch.tensor(rng.uniform(0.0, 1.0, grid.shape).astype(np.float32), device="cuda")
    gathered = module.adjoint_gather(seeds, packed_pose=packed)
    assert torch.equal(gathered, module.adjoint_gather(seeds, packed_pose=packed))
    image = module(values, packed_pose=packed)
    left = math.fsum((image.double() * seeds).reshape(-1).tolist())
    right = math.fsum((values.double() * gathered).reshape(-1).tolist())
    assert abs(left - right) <= 1e-10 * max(abs(left), abs(right))
    field = values.clone().requires_grad_(True)
    (scatter,) = torch.autograd.grad(module(field, packed_pose=packed), field, seeds)
    g, s_ = gathered.cpu().numpy(), scatter.double().cpu().numpy()
    np.testing.assert_allclose(s_, g, rtol=1e-5, atol=1e-6 * np.abs(g).max())
