"""Volume-VJP increment rules on the GPU, through the C ABI.

The reference scatters RN32(factor * weight) per sample and corner, flushes
increments below 2^-126 to signed zeros on the GPU, sums them per cell and
guards the sums against overflow (kernels/projection_fp32.py:240-252,
437-548); its FP64 destination (volume_gradient_precision="float64",
python/dpt/projection.py:871-1018) keeps unflushed FP32 increments
(FP32 sampling) or unrounded FP64 ones (FP64 sampling).  The cases are those
of tests/python/warp_cpu/test_projection_shared_face.py:91-153 and
test_projection_twenty_pair_oracle.py:17-89: cell walks along and across the
axes with seeds x1 and x1000 (factors beyond 1), seeds of 1e-33 (factors
below 2^-102: the per-increment form), +-1e-35 (every increment flushed:
exact zeros), +-2e38 accumulated onto -+3e38 (finite +-1e38), and twenty
FP64 transpose pairs against an independent 47-midpoint matrix at 2e-12.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests import scenes

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt import _native  # noqa: E402

from .test_gpu_parity import NativeScene  # noqa: E402


def _volume(gpu: NativeScene, values, seeds, *, joint=False, fp64=False, initial=None, pose=None):
    """grad_field from dpt_vjp (volume, optionally joint) into an FP32 or FP64
    destination, optionally accumulated onto `initial`."""
    field, _ = gpu._field(values)
    poses = gpu._poses(pose)
    dtype = torch.float32 if gpu.precision == "float32" else torch.float64
    seed = torch.as_tensor(np.ascontiguousarray(seeds).reshape(gpu.shape)).to(dtype).cuda()
    out_dtype = torch.float64 if fp64 else torch.float32
    flags = _native.VOLUME | (_native.VOLUME_FP64 if fp64 else 0)
    if initial is None:
        gv = torch.full((gpu.grid.voxels,), 5.0, dtype=out_dtype, device="cuda")
    else:
        gv = torch.as_tensor(np.ascontiguousarray(initial)).to(out_dtype).cuda()
        flags |= _native.ACCUMULATE
    gp = None
    if joint:
        flags |= _native.POSE
        gp = torch.empty((gpu.shape[0], 12), dtype=torch.float64, device="cuda")
    gpu.plan.vjp(field.data_ptr(), poses.data_ptr(), seed.data_ptr(), gv.data_ptr(),
                 gp.data_ptr() if gp is not None else None, flags, gpu.stream)
    status = gpu.plan.status(gpu.stream)
    return gv.cpu().numpy(), status


# "1": the volume adjoint by slab chunks (vol32_slab_kernel), forced on
SLABS = ["0", "1"]


@pytest.mark.parametrize("slabs", SLABS)
@pytest.mark.parametrize("joint", [False, True])
@pytest.mark.parametrize("seed_scale", [1.0, 1000.0, 1e-33])
@pytest.mark.parametrize("samples", [2, 7, 1024])
@pytest.mark.parametrize("direction", scenes.SHARED_FACE_DIRECTIONS)
def test_shared_face_walks_match_the_per_sample_scatter(direction, samples, seed_scale, joint, slabs, monkeypatch):
    print(f"case shared face {direction} S={samples} scale={seed_scale} joint={joint} slabs={slabs}", flush=True)
    monkeypatch.setenv("DPT_AMD_SLAB_CHUNKS", slabs)
    grid, sources, endpoints, values = scenes.shared_face_scene(direction)
    seeds = (seed_scale * np.asarray([0.7, -0.3, 0.0])).astype(np.float32)
    ref = O.Scene(grid, samples, rays=(sources, endpoints))
    expected, _, _, _ = ref.vjp(values, seeds, pose=False)
    gpu = NativeScene(grid, samples, rays=(sources, endpoints))
    actual, status = _volume(gpu, values, seeds, joint=joint)
    assert status.tolist() == [0]
    assert np.count_nonzero(expected) > 0
    np.testing.assert_allclose(actual, expected, rtol=2e-5, atol=3e-7 * seed_scale)


@pytest.mark.parametrize("slabs", SLABS)
@pytest.mark.parametrize("joint", [False, True])
@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_subnormal_increments_give_exact_zeros(sign, joint, slabs, monkeypatch):
    print(f"case subnormal sign={sign} joint={joint} slabs={slabs}", flush=True)
    monkeypatch.setenv("DPT_AMD_SLAB_CHUNKS", slabs)
    grid, source, target, values = scenes.subnormal_scene()
    seeds = np.asarray([sign * 1e-35], dtype=np.float32)
    gpu = NativeScene(grid, 1024, rays=(source, target))
    actual, status = _volume(gpu, values, seeds, joint=joint)
    assert status.tolist() == [0]
    np.testing.assert_array_equal(actual, np.zeros_like(actual))
    # the FP64 destination keeps the unflushed increments
    unflushed = O.explicit_fp32_scatter(grid, list(zip(source, target)), scenes.IDENTITY12, 1024, seeds, flush=False)
    wide, status = _volume(gpu, values, seeds, joint=joint, fp64=True)
    assert status.tolist() == [0]
    assert np.count_nonzero(unflushed) > 0
    np.testing.assert_allclose(wide, unflushed, rtol=2e-5, atol=0)


@pytest.mark.parametrize("slabs", SLABS)
@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_overflow_scale_seeds_accumulate_finitely_onto_an_opposing_destination(sign, reverse, slabs, monkeypatch):
    print(f"case overflow sign={sign} reverse={reverse} slabs={slabs}", flush=True)
    monkeypatch.setenv("DPT_AMD_SLAB_CHUNKS", slabs)
    grid, source, target, values = scenes.overflow_scene(reverse)
    seeds = np.asarray([sign * 2e38], dtype=np.float32)
    expected, _, _, _ = O.Scene(grid, 1024, rays=(source, target)).vjp(values, seeds, pose=False)
    initial = np.full(grid.voxels, -sign * 3e38, dtype=np.float32)
    gpu = NativeScene(grid, 1024, rays=(source, target))
    actual, status = _volume(gpu, values, seeds, initial=initial)
    assert status.tolist() == [0]
    assert np.all(np.isfinite(actual))
    np.testing.assert_allclose(actual, expected + initial.astype(np.float64), rtol=2e-5, atol=0)
    np.testing.assert_allclose(actual[1], sign * 1e38, rtol=2e-5)


@pytest.mark.parametrize("direction", [(1, 0, 0), (0, -1, 0), (-1, -1, -1)])
@pytest.mark.parametrize("slabs", SLABS)
def test_fp32_sampling_into_an_fp64_destination(direction, slabs, monkeypatch):
    """Unflushed RN32 increments summed in FP64 (the FP64-destination scatter)."""
    print(f"case fp64 destination {direction} slabs={slabs}", flush=True)
    monkeypatch.setenv("DPT_AMD_SLAB_CHUNKS", slabs)
    grid, sources, endpoints, values = scenes.shared_face_scene(direction)
    seeds = np.asarray([0.7, -0.3, 0.0], dtype=np.float32)
    expected = O.explicit_fp32_scatter(grid, list(zip(sources, endpoints)), scenes.IDENTITY12, 97, seeds, flush=False)
    gpu = NativeScene(grid, 97, rays=(sources, endpoints))
    actual, status = _volume(gpu, values, seeds, fp64=True)
    assert status.tolist() == [0]
    assert actual.dtype == np.float64
    np.testing.assert_allclose(actual, expected, rtol=2e-6, atol=1e-12)
    previous = np.random.default_rng(3).uniform(-1, 1, grid.voxels)
    summed, _ = _volume(gpu, values, seeds, fp64=True, initial=previous)
    np.testing.assert_allclose(summed, expected + previous, rtol=0, atol=2e-6 * np.abs(expected).max())
    np.testing.assert_allclose(summed - previous, actual, rtol=0, atol=1e-15)


def test_twenty_fp64_transpose_pairs_against_the_47_midpoint_matrix():
    """tests/python/warp_cpu/test_projection_twenty_pair_oracle.py:17-89
    (midpoint): FP64 sampling, FP64 destination, 2e-12."""
    grid = O.Grid((3, 4, 5), (1.2, 0.8, 1.1), (-2.0, -1.0, -1.0))
    panel = O.Panel((-12.0, 0.31, 0.43), (12.0, -0.6, -0.6), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.071, 0.11), (3, 5))
    pose = scenes.compose(scenes.IDENTITY12, (0.13, -0.08, 0.07, 0.02, -0.03, 0.01))
    rows, cols = panel.shape
    matrix = np.empty((rows * cols, grid.voxels))
    for voxel in range(grid.voxels):
        basis = np.zeros(grid.voxels)
        basis[voxel] = 1.0
        for pixel in range(rows * cols):
            r, c = divmod(pixel, cols)
            matrix[pixel, voxel] = O.integrate_segment(basis, grid, pose, panel.source, panel.pixel(r, c),
                                                       midpoint_samples=47)
    gpu = NativeScene(grid, 47, panels=[panel], precision="float64")
    rng = np.random.default_rng(41072)
    for pair in range(20):
        values = np.abs(rng.normal(size=grid.voxels)).astype(np.float32)
        seeds = rng.normal(size=rows * cols)
        image, status = gpu.forward(values, pose)
        assert status.tolist() == [0]
        np.testing.assert_allclose(image.ravel(), matrix @ values.astype(np.float64), rtol=2e-12, atol=2e-12,
                                   err_msg=f"forward pair {pair}")
        gradient, status = _volume(gpu, values, seeds, fp64=True, pose=pose)
        assert status.tolist() == [0]
        np.testing.assert_allclose(gradient, matrix.T @ seeds, rtol=2e-12, atol=2e-12, err_msg=f"adjoint pair {pair}")
        left, right = float(image.ravel() @ seeds), float(values.astype(np.float64) @ gradient)
        assert abs(left - right) < 2e-12 * max(1.0, abs(left), abs(right))
