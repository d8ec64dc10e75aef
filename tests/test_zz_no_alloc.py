"""The allocation-free adjoint gather (dpt_adjoint_gather_noalloc) on the GPU.

It runs the kernels of dpt_adjoint_gather on the library's fixed per-device
workspace instead of a plan-owned ray table, so its output must equal
dpt_adjoint_gather's bit for bit (FP64 and FP32 destinations, with and
without accumulation).  Two calls in a row, the second on a smaller plan,
check that the shared workspace keeps nothing of the previous call that
matters (the guard word is cleared and both tables are rewritten by every
call).  A plan whose ray table exceeds the workspace is refused before any
launch.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests import scenes

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt import _native  # noqa: E402
from dpt.contracts import ContractError  # noqa: E402

from .test_gpu_parity import NativeScene  # noqa: E402


def _run(gpu: NativeScene, seeds, pose, *, noalloc: bool, precision: str, previous=None):
    poses = gpu._poses(pose)
    cot = torch.as_tensor(np.ascontiguousarray(seeds, dtype=np.float64).reshape(gpu.shape)).cuda()
    dtype = torch.float64 if precision == "float64" else torch.float32
    flags = 0
    if previous is None:
        out = torch.full((gpu.grid.voxels,), -7.0, dtype=dtype, device="cuda")
    else:
        out = torch.as_tensor(np.ascontiguousarray(previous)).to(dtype).cuda()
        flags |= _native.ACCUMULATE
    call = gpu.plan.adjoint_gather_noalloc if noalloc else gpu.plan.adjoint_gather
    call(poses.data_ptr(), cot.data_ptr(), out.data_ptr(), precision, flags, gpu.stream)
    status = gpu.plan.status(gpu.stream)
    return out.cpu().numpy(), status


def _large():
    grid = O.Grid((32, 30, 28), (1.0, 1.1, 0.9), (-15.0, -16.0, -12.0))
    panels = scenes.cone_panels(grid, 4, 64, 72, sod=90.0, sdd=150.0, offset=0.25)
    pose = scenes.compose(scenes.IDENTITY12, (0.4, -0.3, 0.2, 0.02, -0.01, 0.015))
    return grid, dict(panels=panels), pose


def _check_pair(grid, geometry, pose, samples, seed):
    print(f"case noalloc gather grid={grid.shape} S={samples} seed={seed}", flush=True)
    gpu = NativeScene(grid, samples, precision="float64", **geometry)
    seeds = np.random.default_rng(seed).uniform(-1.0, 1.0, int(np.prod(gpu.shape)))
    seeds[::5] = 0.0
    for precision in ("float64", "float32"):
        expected, status_e = _run(gpu, seeds, pose, noalloc=False, precision=precision)
        actual, status_a = _run(gpu, seeds, pose, noalloc=True, precision=precision)
        assert status_e.tolist() == status_a.tolist() == [0] * gpu.shape[0]
        assert np.count_nonzero(expected) > 0
        assert actual.tobytes() == expected.tobytes()
    previous = np.random.default_rng(seed + 1).uniform(-1.0, 1.0, grid.voxels)
    expected, _ = _run(gpu, seeds, pose, noalloc=False, precision="float64", previous=previous)
    actual, status = _run(gpu, seeds, pose, noalloc=True, precision="float64", previous=previous)
    assert status.tolist() == [0] * gpu.shape[0]
    assert actual.tobytes() == expected.tobytes()
    return gpu, seeds


def test_noalloc_gather_equals_gather_on_two_calls_in_a_row():
    # first call: 4 views of 64 x 72 over a 32 x 30 x 28 grid
    grid, geometry, pose = _large()
    _check_pair(grid, geometry, pose, 64, 31)
    # second call, smaller: the six-ray bundle with missing and axis-parallel
    # rays; also against the oracle's gather (bitwise)
    grid, geometry, pose = scenes.gather_scene("bundle")
    gpu, seeds = _check_pair(grid, geometry, pose, 31, 32)
    expected, _ = O.Scene(grid, 31, **geometry).gather(seeds, pose)
    actual, _ = _run(gpu, seeds, pose, noalloc=True, precision="float64")
    np.testing.assert_array_equal(actual.view(np.int64), expected.view(np.int64))


def test_noalloc_gather_refuses_plans_beyond_its_workspace():
    # two 512 x 512 views need 2 * 262144 * 112 bytes of ray table > 32 MiB
    grid = O.Grid((8, 8, 8), (1.0, 1.0, 1.0), (-3.5, -3.5, -3.5))
    panels = scenes.cone_panels(grid, 2, 512, 512, sod=60.0, sdd=100.0)
    gpu = NativeScene(grid, 8, panels=panels, precision="float64")
    assert 112 * 2 * 512 * 512 > _native.GATHER_NOALLOC_BYTES
    poses = gpu._poses(None)
    cot = torch.zeros(gpu.shape, dtype=torch.float64, device="cuda")
    out = torch.zeros(grid.voxels, dtype=torch.float64, device="cuda")
    with pytest.raises(ContractError, match="fixed workspace"):
        gpu.plan.adjoint_gather_noalloc(poses.data_ptr(), cot.data_ptr(), out.data_ptr(), "float64", 0, gpu.stream)
