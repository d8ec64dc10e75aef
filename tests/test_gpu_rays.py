"""Ray-coordinate and cone-intrinsic derivatives on the GPU.

dpt_ray_vjp returns, per ray, cotangent x d(optical depth)/d(world source,
world endpoint) through the plan's arithmetic and rule; the oracle's
ray-coordinate mode (pinned in test_oracle.py by central differences) is the
checker: 1e-5 relative L2 on the FP32 path, 1e-10 on FP64 midpoint and
cell-Gauss.  Two consistency checks need no oracle: moving the object by dt
moves every segment by -dt, so the packed pose gradient's translation part
is minus the per-view sum of source and endpoint derivatives; and the cone
intrinsic VJP (python/dpt/batched_projection.py:340-539) matches central
differences of the FP64 forward in SOD, SDD and the detector offsets.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import scenes

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt import _native  # noqa: E402
from dpt.geometry import ConeBeamIntrinsics  # noqa: E402
from dpt.torch import Projector  # noqa: E402
from dpt.volumes import GridSpec  # noqa: E402

from .test_gpu_gather import _case  # noqa: E402
from .test_gpu_parity import NativeScene  # noqa: E402

MODES = [("float32", "midpoint"), ("float64", "midpoint"), ("float64", "cell_gauss")]


def _ray_vjp(gpu: NativeScene, values, seeds, pose):
    field, _ = gpu._field(values)
    poses = gpu._poses(pose)
    dtype = torch.float32 if gpu.precision == "float32" else torch.float64
    seed = torch.as_tensor(np.ascontiguousarray(seeds).reshape(gpu.shape)).to(dtype).cuda()
    out = torch.empty((*gpu.shape, 6), dtype=torch.float64, device="cuda")
    gpu.plan.ray_vjp(field.data_ptr(), poses.data_ptr(), seed.data_ptr(), out.data_ptr(), 0, gpu.stream)
    status = gpu.plan.status(gpu.stream)
    return out.cpu().numpy().reshape(-1, 6), status


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("mode", MODES, ids=lambda m: f"{m[0]}-{m[1]}")
@pytest.mark.parametrize("name", ["oblique", "cone", "thin_x"])
def test_ray_vjp_matches_the_oracle(name, mode):
    precision, integration = mode
    print(f"case ray vjp {name} {precision} {integration}", flush=True)
    grid, geometry, pose = _case(name)
    extra = _native.CELL_GAUSS if integration == "cell_gauss" else 0
    gpu = NativeScene(grid, 97, precision=precision, extra_needs=extra, **geometry)
    ref = O.Scene(grid, 97, **geometry)
    values = scenes.smooth_field(grid, 4).reshape(-1)
    seeds = np.random.default_rng(5).uniform(-1.0, 1.0, int(np.prod(gpu.shape)))
    arith = "fp32" if precision == "float32" else "fp64"
    expected, ref_status = ref.ray_vjp(values, seeds, pose, precision=arith, integration=integration)
    actual, status = _ray_vjp(gpu, values, seeds, pose)
    assert status.tolist() == ref_status.tolist() == [0] * gpu.shape[0]
    assert _rel(actual, expected) <= (1e-5 if precision == "float32" else 1e-10)
    # translation part of the packed pose gradient = -(sum of source + endpoint derivatives)
    _, packed, _ = gpu.vjp(values, seeds, pose, volume=False)
    rays = actual.reshape(gpu.shape[0], -1, 6)
    translation = -(rays[:, :, :3] + rays[:, :, 3:]).sum(axis=1)
    assert _rel(packed[:, 9:], translation) <= (2e-5 if precision == "float32" else 1e-10)


def test_cone_intrinsics_vjp_matches_central_differences():
    grid = GridSpec((20, 21, 22), (1.0, 1.0, 1.0), (-10.5, -10.0, -9.5))
    values = torch.tensor(scenes.smooth_field(O.Grid(grid.shape, grid.spacing_mm, grid.origin_mm), 9), device="cuda")
    angles = (0.1, 1.3, 2.6)
    base = dict(source_object_distance_mm=80.0, source_detector_distance_mm=130.0, shape=(14, 16),
                spacing_mm=(2.4, 2.4), detector_offset_mm=(1.5, -2.0))

    def module(**change):
        intrinsics = ConeBeamIntrinsics(**{**base, **change})
        return intrinsics, Projector(grid=grid, geometry=intrinsics.batch(angles), samples_per_ray=256,
                                     precision="fp64")

    intrinsics, proj = module()
    seed = torch.linspace(-1.0, 1.0, 3 * 14 * 16, dtype=torch.float64, device="cuda").reshape(3, 14, 16)
    grad = proj.intrinsics_vjp(values, seed, intrinsics, angles).sum(dim=0).cpu().numpy()

    def loss(**change):
        _, m = module(**change)
        with torch.no_grad():
            return float((m(values) * seed).sum())

    h = 1e-3
    fd = [
        (loss(source_object_distance_mm=80.0 + h) - loss(source_object_distance_mm=80.0 - h)) / (2 * h),
        (loss(source_detector_distance_mm=130.0 + h) - loss(source_detector_distance_mm=130.0 - h)) / (2 * h),
        (loss(detector_offset_mm=(1.5 + h, -2.0)) - loss(detector_offset_mm=(1.5 - h, -2.0))) / (2 * h),
        (loss(detector_offset_mm=(1.5, -2.0 + h)) - loss(detector_offset_mm=(1.5, -2.0 - h))) / (2 * h),
    ]
    scale = max(abs(x) for x in fd)
    for k in range(4):
        assert math.isclose(grad[k], fd[k], rel_tol=2e-3, abs_tol=1e-4 * scale), (k, grad[k], fd[k])
    with pytest.raises(Exception, match="intrinsics"):
        proj.intrinsics_vjp(values, seed, ConeBeamIntrinsics(**{**base, "source_object_distance_mm": 81.0}), angles)


@pytest.mark.parametrize("mode", MODES, ids=lambda m: f"{m[0]}-{m[1]}")
def test_projector_ray_vjp_matches_the_oracle(mode):
    """Projector.ray_vjp through the host layer (a plan made without pose
    buffers), against the oracle's ray-coordinate mode."""
    from dpt.geometry import DetectorGeometry

    precision, integration = mode
    grid = O.Grid((10, 11, 12), (1.0, 1.0, 1.0), (-5.5, -5.0, -4.5))
    panels = scenes.cone_panels(grid, 2, 8, 9, sod=50.0, sdd=90.0, offset=0.2)
    geometry = [DetectorGeometry(p.source, p.origin, p.u, p.v, p.spacing, p.shape) for p in panels]
    module = Projector(grid=GridSpec(grid.shape, grid.spacing, grid.origin), geometry=geometry, samples_per_ray=64,
                       precision="fp32" if precision == "float32" else "fp64", integration=integration)
    values = scenes.smooth_field(grid, 2)
    dtype = torch.float32 if precision == "float32" else torch.float64
    seed = torch.linspace(-1.0, 1.0, 2 * 8 * 9, dtype=dtype, device="cuda").reshape(2, 8, 9)
    rays = module.ray_vjp(torch.tensor(values, device="cuda"), seed)
    assert tuple(rays.shape) == (2, 8, 9, 6)
    expected, _ = O.Scene(grid, 64, panels=panels).ray_vjp(
        values, seed.cpu().numpy(), precision="fp32" if precision == "float32" else "fp64", integration=integration)
    assert _rel(rays.cpu().numpy().reshape(-1, 6), expected) <= (1e-5 if precision == "float32" else 1e-10)


@pytest.mark.parametrize("storage", ["bfloat16", "float16"])
def test_ray_vjp_with_half_storage_matches_the_oracle_on_the_rounded_field(storage):
    """Half storage rounds the field once; the ray derivatives of that field
    (FP32 sampling, compensated sums) match the oracle's FP32 path on the
    rounded values."""
    from dpt.geometry import DetectorGeometry

    grid = O.Grid((10, 11, 12), (1.0, 1.0, 1.0), (-5.5, -5.0, -4.5))
    panels = scenes.cone_panels(grid, 2, 8, 9, sod=50.0, sdd=90.0, offset=0.2)
    geometry = [DetectorGeometry(p.source, p.origin, p.u, p.v, p.spacing, p.shape) for p in panels]
    module = Projector(grid=GridSpec(grid.shape, grid.spacing, grid.origin), geometry=geometry, samples_per_ray=64,
                       storage=storage)
    values = torch.tensor(scenes.smooth_field(grid, 2), device="cuda")
    rounded = values.to(torch.bfloat16 if storage == "bfloat16" else torch.float16).float()
    seed = torch.linspace(-1.0, 1.0, 2 * 8 * 9, dtype=torch.float32, device="cuda").reshape(2, 8, 9)
    rays = module.ray_vjp(values, seed)
    expected, _ = O.Scene(grid, 64, panels=panels).ray_vjp(rounded.cpu().numpy(), seed.cpu().numpy(), precision="fp32")
    assert _rel(rays.cpu().numpy().reshape(-1, 6), expected) <= 1e-5
