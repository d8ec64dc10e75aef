"""Cell-layout gathers on the GPU.

The cell layout stores each interpolation cell's 8 corner values
contiguously, so a trilinear gather is one or two 16-byte loads.  It changes
where the corner values come from, never their values or the arithmetic, so
the bar is bit-identity with the linear-layout launches (images and pose
gradients; volume gradients are FP64 atomic sums rounded to FP32 and are
compared to 1e-7), plus the linear-layout parity with the oracle that
test_gpu_parity.py holds.  The relayout itself is checked bit for bit against
a numpy restatement of the layout on grids with singleton axes.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests import scenes

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt import _native  # noqa: E402

from .test_gpu_parity import NativeScene, _cone  # noqa: E402


def _numpy_cells(values: np.ndarray) -> np.ndarray:
    """(nz, ny, nx) -> (cells, 8) with corner q = (x + (q&1), y + (q>>1&1), z + (q>>2&1)), clamped."""
    nz, ny, nx = values.shape
    mz, my, mx = max(nz - 1, 1), max(ny - 1, 1), max(nx - 1, 1)
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    corners = []
    for q in range(8):
        x = np.minimum(i + (q & 1), nx - 1)
        y = np.minimum(j + ((q >> 1) & 1), ny - 1)
        z = np.minimum(k + ((q >> 2) & 1), nz - 1)
        corners.append(values[z, y, x].reshape(-1))
    return np.stack(corners, axis=1)


@pytest.mark.parametrize("storage", ["float32", "bfloat16", "float16"])
@pytest.mark.parametrize("shape", [(5, 6, 7), (1, 4, 9), (6, 1, 3), (4, 5, 1), (1, 1, 1)])
def test_convert_cells_is_the_layout_bit_for_bit(storage, shape):
    print(f"case cells {storage} {shape}", flush=True)
    rng = np.random.default_rng(11)
    dtype = {"float32": torch.float32, "bfloat16": torch.bfloat16, "float16": torch.float16}[storage]
    src = torch.as_tensor(rng.uniform(0.0, 2.0, shape).astype(np.float32)).to(dtype).cuda()
    count = _native.cell_count(shape)
    assert count == int(np.prod([max(n - 1, 1) for n in shape]))
    dst = torch.full((count * 8,), -1.0, dtype=dtype, device="cuda")
    _native.convert_cells(src.data_ptr(), dst.data_ptr(), shape, storage, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    bits = torch.int32 if storage == "float32" else torch.int16
    expected = _numpy_cells(src.view(bits).cpu().numpy())
    np.testing.assert_array_equal(dst.view(bits).cpu().numpy().reshape(count, 8), expected)


def _pair(monkeypatch, grid, samples, storage, **geometry):
    monkeypatch.setenv("DPT_AMD_GATHER", "linear")
    linear = NativeScene(grid, samples, storage=storage, **geometry)
    monkeypatch.setenv("DPT_AMD_GATHER", "cells")
    cells = NativeScene(grid, samples, storage=storage, **geometry)
    return linear, cells


def _thin_scene(axis):
    shape = [7, 8, 9]
    shape[axis] = 1
    grid = O.Grid(tuple(shape), (1.0, 0.9, 1.1), (-4.0, -3.5, -3.0))
    panels = scenes.cone_panels(grid, 3, 12, 14, sod=60.0, sdd=100.0, offset=0.2)
    return grid, panels, scenes.smooth_field(grid, 4)


CASES = ["oblique", "workflow", "cone", "thin_z", "thin_x"]


def _case(name):
    if name == "oblique":
        grid, src, end, pose, values = scenes.oblique_scene()
        return grid, values, pose, dict(rays=(src, end))
    if name == "workflow":
        grid, panels, values = scenes.workflow_scene()
        return grid, values, scenes.compose(scenes.IDENTITY12, (0.03, -0.02, 0.01, 0.04, -0.03, 0.02)), dict(
            panels=panels)
    if name == "cone":
        grid, panels, values = _cone()
    else:
        grid, panels, values = _thin_scene(0 if name == "thin_z" else 2)
    return grid, values, scenes.compose(scenes.IDENTITY12, (0.5, -0.7, 0.3, 0.02, -0.015, 0.03)), dict(
        panels=panels)


@pytest.mark.parametrize("storage", ["float32", "bfloat16", "float16"])
@pytest.mark.parametrize("name", CASES)
def test_cell_gathers_equal_linear_gathers(name, storage, monkeypatch):
    print(f"case cell gather {name} {storage}", flush=True)
    grid, values, pose, geometry = _case(name)
    linear, cells = _pair(monkeypatch, grid, 97, storage, **geometry)
    a, sa = linear.forward(values, pose)
    b, sb = cells.forward(values, pose)
    assert sa.tolist() == sb.tolist() == [0] * len(sa)
    np.testing.assert_array_equal(b, a)
    seeds = np.random.default_rng(6).uniform(-1.0, 1.0, a.size).astype(np.float32)
    for components in (12, 6):
        ga, pa, _ = linear.vjp(values, seeds, pose, components=components)
        gb, pb, status = cells.vjp(values, seeds, pose, components=components)
        assert status.tolist() == [0] * len(sa)
        np.testing.assert_array_equal(pb, pa)
        np.testing.assert_allclose(gb, ga, rtol=1e-7, atol=1e-12)
    _, pa, _ = linear.vjp(values, seeds, pose, volume=False)
    _, pb, _ = cells.vjp(values, seeds, pose, volume=False)
    np.testing.assert_array_equal(pb, pa)


def test_cell_gathers_match_the_oracle(monkeypatch):
    monkeypatch.setenv("DPT_AMD_GATHER", "cells")
    grid, values, pose, geometry = _case("cone")
    ref = O.Scene(grid, 97, panels=geometry["panels"])
    gpu = NativeScene(grid, 97, **geometry)
    expected, _ = ref.forward(values, pose)
    actual, status = gpu.forward(values, pose)
    assert status.tolist() == [0] * len(status)
    assert O.relative_l2(actual, expected) <= 1e-6
    seeds = np.random.default_rng(2).uniform(-1, 1, expected.size).astype(np.float32)
    g_ref, p_ref, _, _ = ref.vjp(values, seeds, pose)
    g_gpu, p_gpu, _ = gpu.vjp(values, seeds, pose)
    assert O.relative_l2(g_gpu, g_ref) <= 1e-5
    assert O.relative_l2(p_gpu, p_ref) <= 1e-5
