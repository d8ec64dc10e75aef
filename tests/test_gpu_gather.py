"""Deterministic adjoint gather on the GPU, through the C ABI.

The gather sums every voxel's addends in the reference's fixed order
(kernels/projection_adjoint.py:73-170) in FP64, so the bar is bit-identity
with the oracle's gather (itself pinned in test_oracle.py against a
voxel-driven restatement and the transpose of the explicit FP64 forward
matrix), in FP64 and FP32 destinations, with and without accumulation, and
bit-identical repeats.  At larger sizes, where the oracle's cost grows, the
checks are size-independent: the adjoint identity with the GPU FP64 forward
(1e-10 relative), agreement with the atomic FP64 volume VJP (1e-6: its
increments are FP32-rounded) and bit-identical repeats.

Edge cases follow tests/python/warp_cpu/test_projection_adjoint_gather.py and
test_projection_culled_adjoint.py: a non-square panel, a bundle with
axis-parallel and missing rays, a clamped 1x2x3 grid, singleton axes, a
source close to the grid (wide pixel windows), a non-finite cotangent and a
non-rigid pose (destination untouched), and an FP32 plan (rejected).
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
from dpt.contracts import ContractError  # noqa: E402

from .test_gpu_parity import NativeScene, _cone  # noqa: E402


def _gather(gpu: NativeScene, seeds, pose, *, precision="float64", previous=None, flags=0):
    poses = gpu._poses(pose)
    cot = torch.as_tensor(np.ascontiguousarray(seeds, dtype=np.float64).reshape(gpu.shape)).cuda()
    dtype = torch.float64 if precision == "float64" else torch.float32
    if previous is None:
        out = torch.full((gpu.grid.voxels,), 19.0, dtype=dtype, device="cuda")
    else:
        out = torch.as_tensor(np.ascontiguousarray(previous)).to(dtype).cuda()
        flags |= _native.ACCUMULATE
    gpu.plan.adjoint_gather(poses.data_ptr(), cot.data_ptr(), out.data_ptr(), precision, flags, gpu.stream)
    status = gpu.plan.status(gpu.stream)
    return out.cpu().numpy(), status


def _thin(axis):
    shape = [7, 8, 9]
    shape[axis] = 1
    grid = O.Grid(tuple(shape), (1.0, 0.9, 1.1), (-4.0, -3.5, -3.0))
    return grid, dict(panels=scenes.cone_panels(grid, 3, 12, 14, sod=60.0, sdd=100.0, offset=0.2))


def _case(name):
    """(grid, geometry kwargs, pose) of each parity case."""
    if name in scenes.GATHER_CASES:
        return scenes.gather_scene(name)
    if name == "workflow":
        grid, panels, _ = scenes.workflow_scene(3)
        return grid, dict(panels=panels), scenes.compose(scenes.IDENTITY12, (0.03, -0.02, 0.01, 0.04, -0.03, 0.02))
    if name == "oblique":
        grid, src, end, pose, _ = scenes.oblique_scene()
        return grid, dict(rays=(src, end)), pose
    if name == "cone":
        grid, panels, _ = _cone()
        return grid, dict(panels=panels), scenes.compose(scenes.IDENTITY12, (0.5, -0.7, 0.3, 0.02, -0.015, 0.03))
    if name == "near_source":
        # source 3 mm from the grid face: wide, strongly varying pixel windows
        grid = O.Grid((10, 11, 12), (1.0, 1.0, 1.0), (-5.5, -5.0, -4.5))
        panels = scenes.cone_panels(grid, 4, 16, 18, sod=9.0, sdd=30.0, offset=0.4)
        return grid, dict(panels=panels), scenes.compose(scenes.IDENTITY12, (0.1, 0.2, -0.1, 0.05, 0.02, -0.03))
    if name in ("thin_z", "thin_x"):
        grid, geometry = _thin(0 if name == "thin_z" else 2)
        return grid, geometry, scenes.compose(scenes.IDENTITY12, (0.1, -0.2, 0.05, 0.02, -0.01, 0.03))
    raise KeyError(name)


CASES = [*scenes.GATHER_CASES, "workflow", "oblique", "cone", "near_source", "thin_z", "thin_x"]


def _pair(name, samples):
    grid, geometry, pose = _case(name)
    return grid, pose, O.Scene(grid, samples, **geometry), NativeScene(grid, samples, precision="float64", **geometry)


@pytest.mark.parametrize("samples", [31, 97])
@pytest.mark.parametrize("name", CASES)
def test_gather_is_bitwise_the_oracle_gather(name, samples):
    print(f"case gather {name} S={samples}", flush=True)
    grid, pose, ref, gpu = _pair(name, samples)
    seeds = np.random.default_rng(584291).uniform(-1.0, 1.0, int(np.prod(gpu.shape)))
    seeds[::7] = 0.0  # zero cotangents skip their rays
    expected, ref_status = ref.gather(seeds, pose)
    actual, status = _gather(gpu, seeds, pose)
    assert status.tolist() == ref_status.tolist() == [0] * gpu.shape[0]
    assert np.count_nonzero(expected) > 0
    np.testing.assert_array_equal(actual.view(np.int64), expected.view(np.int64))
    single, status = _gather(gpu, seeds, pose, precision="float32")
    assert status.tolist() == [0] * gpu.shape[0]
    np.testing.assert_array_equal(single.view(np.int32), expected.astype(np.float32).view(np.int32))


@pytest.mark.parametrize("name", ["cone", "bundle", "thin_x"])
def test_gather_repeats_and_accumulates_bit_for_bit(name):
    print(f"case gather repeat {name}", flush=True)
    grid, pose, ref, gpu = _pair(name, 64)
    seeds = np.random.default_rng(9).uniform(-1.0, 1.0, int(np.prod(gpu.shape)))
    first, _ = _gather(gpu, seeds, pose)
    again, _ = _gather(gpu, seeds, pose)
    assert first.tobytes() == again.tobytes()
    previous = np.random.default_rng(10).uniform(-1.0, 1.0, grid.voxels)
    expected, _ = ref.gather(seeds, pose, previous=previous)
    actual, status = _gather(gpu, seeds, pose, previous=previous)
    assert status.tolist() == [0] * gpu.shape[0]
    np.testing.assert_array_equal(actual.view(np.int64), expected.view(np.int64))
    previous32 = previous.astype(np.float32)
    expected32, _ = ref.gather(seeds, pose, previous=previous32, precision="float32")
    actual32, _ = _gather(gpu, seeds, pose, previous=previous32, precision="float32")
    np.testing.assert_array_equal(actual32.view(np.int32), expected32.view(np.int32))


def test_gather_rejects_bad_inputs_without_writing():
    """tests/python/warp_cpu/test_projection_adjoint_gather.py:198-235."""
    grid, pose, _, gpu = _pair("clamped", 31)
    seeds = np.ones(int(np.prod(gpu.shape)))
    seeds[1] = math.nan
    out, status = _gather(gpu, seeds, pose)
    assert status.tolist() == [2]
    assert out.tolist() == [19.0] * grid.voxels
    scaled = list(scenes.IDENTITY12)
    scaled[0] = 1.01
    out, status = _gather(gpu, np.ones_like(seeds), scaled)
    assert status.tolist() == [1]
    assert out.tolist() == [19.0] * grid.voxels
    out, status = _gather(gpu, np.ones_like(seeds), pose)  # statuses were consumed
    assert status.tolist() == [0] and np.isfinite(out).all()
    # overflow of the accumulated value is reported (1e308 + 1e308)
    tiny = O.Grid((1, 1, 1), (1.0, 1.0, 1.0))
    one = NativeScene(tiny, 3, rays=(np.array([[0.0, 0.0, -2.0]]), np.array([[0.0, 0.0, 2.0]])), precision="float64")
    out, status = _gather(one, [1e308], None, previous=np.array([1e308]))
    assert status.tolist() == [2] and not np.isfinite(out).all()
    fp32 = NativeScene(grid, 31, rays=scenes.gather_scene("clamped")[1]["rays"])
    with pytest.raises(ContractError, match="float64"):
        _gather(fp32, np.ones_like(seeds), pose)


def test_gather_matches_adjoint_identity_and_scatter_at_size():
    """128^3 grid, 24 views of 96x112 pixels, 256 samples: properties that do
    not need the oracle."""
    n = 128
    grid = O.Grid((n, n, n), (1.0, 1.0, 1.0), (-63.5, -63.5, -63.5))
    panels = scenes.cone_panels(grid, 24, 96, 112, sod=400.0, sdd=700.0, offset=0.1)
    pose = scenes.compose(scenes.IDENTITY12, (0.8, -0.5, 0.3, 0.02, -0.01, 0.015))
    print("case gather 128^3 x 24 x 96x112 S=256", flush=True)
    gpu = NativeScene(grid, 256, panels=panels, precision="float64")
    rng = np.random.default_rng(21)
    seeds = rng.uniform(-1.0, 1.0, int(np.prod(gpu.shape)))
    values = rng.uniform(0.0, 1.0, grid.voxels).astype(np.float32)
    gathered, status = _gather(gpu, seeds, pose)
    assert status.tolist() == [0] * 24
    again, _ = _gather(gpu, seeds, pose)
    assert gathered.tobytes() == again.tobytes()
    image, status = gpu.forward(values, pose)
    assert status.tolist() == [0] * 24
    left = math.fsum(image.reshape(-1) * seeds)
    right = math.fsum(values.astype(np.float64) * gathered)
    assert abs(left - right) <= 1e-10 * max(abs(left), abs(right))
    scatter, _, status = gpu.vjp(values, seeds, pose, pose_grad=False)
    assert status.tolist() == [0] * 24
    scale = np.abs(gathered).max()
    np.testing.assert_allclose(scatter, gathered, rtol=1e-5, atol=1e-6 * scale)


def test_projector_adjoint_gather_matches_the_oracle_and_the_module_adjoint():
    """Projector.adjoint_gather: FP64 and FP32 destinations, accumulation,
    and the contract errors of projection_adjoint_gather."""
    from dpt import phantoms
    from dpt.torch import Projector

    grid = phantoms.centred_grid((14, 15, 16))
    geometry, _ = phantoms.circular_geometry(grid, 3, (12, 13), sod=60.0, sdd=100.0)
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=40, precision="fp64")
    xi = torch.tensor([0.2, -0.1, 0.3, 0.01, 0.02, -0.015], dtype=torch.float64, device="cuda")
    gen = torch.Generator("cuda").manual_seed(4)
    c = torch.randn(module.image_shape, dtype=torch.float64, device="cuda", generator=gen)
    out = module.adjoint_gather(c, xi)
    assert out.shape == grid.shape and out.dtype == torch.float64
    poses = module._poses(c, xi, None)
    ref = O.Scene(
        O.Grid(grid.shape, grid.spacing_mm, grid.origin_mm, grid.orientation),
        40,
        panels=[O.Panel(p.source_mm, p.origin_mm, p.u, p.v, p.spacing_mm, p.shape) for p in geometry.views],
    )
    expected, _ = ref.gather(c.cpu().numpy(), poses.cpu().numpy())
    np.testing.assert_array_equal(out.cpu().numpy().reshape(-1).view(np.int64), expected.view(np.int64))
    single = module.adjoint_gather(c, xi, volume_gradient_precision="float32")
    np.testing.assert_array_equal(single.cpu().numpy().reshape(-1), expected.astype(np.float32))
    module.adjoint_gather(c, xi, out=out, accumulate=True)
    np.testing.assert_array_equal(out.cpu().numpy().reshape(-1), 2.0 * expected)
    # same transpose as the module's autograd adjoint (atomic, FP32-rounded increments)
    field = torch.zeros(grid.shape, dtype=torch.float32, device="cuda", requires_grad=True)
    (adjoint,) = torch.autograd.grad(module(field, xi), field, c)
    scale = float(np.abs(expected).max())
    np.testing.assert_allclose(adjoint.cpu().numpy().reshape(-1), expected, rtol=1e-5, atol=1e-6 * scale)
    with pytest.raises(ContractError, match="float64"):
        Projector(grid=grid, geometry=geometry, samples_per_ray=40).adjoint_gather(c.float(), xi)
    with pytest.raises(ContractError, match="cotangent"):
        module.adjoint_gather(c.float(), xi)
    with pytest.raises(ContractError, match="accumulate"):
        module.adjoint_gather(c, xi, accumulate=True)
