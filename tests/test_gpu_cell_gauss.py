"""Exact cell-Gauss integration on the GPU (FP64 plans with DPT_CELL_GAUSS).

The kernels restate the oracle's cell walk operation for operation
(kernels/projection.py:522-746), so images are compared bit for bit; volume
gradients are FP32-rounded increments summed in FP64 in another order
(2e-7), pose gradients per-ray FP64 terms reduced in another order (1e-10).
The oracle's cell-Gauss rule is pinned in test_oracle.py against the
independent exact piecewise-Gauss integral, closed forms, the adjoint
identity and central differences.  Cases: an oriented grid with axis-aligned
and missing rays, cone panels, thin grids, the clamped 1x2x3 grid and the
six-ray bundle of the gather fixtures.
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

from .test_gpu_gather import _case  # noqa: E402
from .test_gpu_parity import NativeScene  # noqa: E402

CASES = ["oblique", "cone", "near_source", "thin_z", "thin_x", "clamped", "bundle", "workflow"]


def _values(grid, seed=3):
    return scenes.smooth_field(grid, seed).reshape(-1)


def _pair(name, samples=7):
    grid, geometry, pose = _case(name)
    ref = O.Scene(grid, samples, **geometry)
    gpu = NativeScene(grid, samples, precision="float64", extra_needs=_native.CELL_GAUSS, **geometry)
    return grid, pose, ref, gpu


@pytest.mark.parametrize("name", CASES)
def test_cell_gauss_forward_is_bitwise_the_oracle(name):
    print(f"case cell-gauss forward {name}", flush=True)
    grid, pose, ref, gpu = _pair(name)
    values = _values(grid)
    expected, ref_status = ref.forward(values, pose, precision="fp64", integration="cell_gauss")
    actual, status = gpu.forward(values, pose)
    assert status.tolist() == ref_status.tolist()
    np.testing.assert_array_equal(actual.view(np.int64), expected.view(np.int64))


@pytest.mark.parametrize("components", [12, 6])
@pytest.mark.parametrize("name", CASES)
def test_cell_gauss_vjp_matches_the_oracle(name, components):
    print(f"case cell-gauss vjp {name} C={components}", flush=True)
    grid, pose, ref, gpu = _pair(name)
    values = _values(grid)
    seeds = np.random.default_rng(8).uniform(-1.0, 1.0, int(np.prod(gpu.shape)))
    g_ref, p_ref, _, ref_status = ref.vjp(values, seeds, pose, precision="fp64", components=components,
                                          integration="cell_gauss")
    g, p, status = gpu.vjp(values, seeds, pose, components=components)
    assert status.tolist() == ref_status.tolist()
    np.testing.assert_allclose(g, g_ref, rtol=2e-7, atol=1e-12 * max(1.0, np.abs(g_ref).max()))
    np.testing.assert_allclose(p, p_ref, rtol=1e-10, atol=1e-12 * max(1.0, np.abs(p_ref).max()))
    g_only, _, _ = gpu.vjp(values, seeds, pose, pose_grad=False, components=components)
    np.testing.assert_allclose(g_only, g_ref, rtol=2e-7, atol=1e-12 * max(1.0, np.abs(g_ref).max()))
    _, p_only, _ = gpu.vjp(values, seeds, pose, volume=False, components=components)
    np.testing.assert_allclose(p_only, p_ref, rtol=1e-10, atol=1e-12 * max(1.0, np.abs(p_ref).max()))


@pytest.mark.parametrize("mode", ["both", "volume"])
@pytest.mark.parametrize("cap,chunks", [(None, None), ("512", "1"), ("512", "64"), (None, "1"), (None, "64"),
                                        ("8192", "3")])
def test_cell_gauss_volume_bricks_match_the_oracle(cap, chunks, mode, monkeypatch):
    """The volume adjoint's lockstep chunks and LDS bricks: the plan's own
    chunking, one chunk per ray, 64 chunks (shorter than a cell), the largest
    brick, and the smallest (512 entries), where chunks overflow, are halved
    and at 1/8 add to the device accumulator."""
    print(f"case cell-gauss bricks cap={cap} chunks={chunks} {mode}", flush=True)
    if cap is not None:
        monkeypatch.setenv("DPT_AMD_BRICK_CAP", cap)
    if chunks is not None:
        monkeypatch.setenv("DPT_AMD_GAUSS_CHUNKS", chunks)
    n = 24
    grid = O.Grid((n, n + 2, n + 4), (1.0, 0.9, 1.1), (-11.0, -12.0, -13.0))
    panels = scenes.cone_panels(grid, 5, 20, 28, sod=120.0, sdd=200.0, offset=0.3)
    values = scenes.smooth_field(grid, 6)
    pose = scenes.compose(scenes.IDENTITY12, (0.5, -0.7, 0.3, 0.02, -0.015, 0.03))
    ref = O.Scene(grid, 7, panels=panels)
    gpu = NativeScene(grid, 7, panels=panels, precision="float64", extra_needs=_native.CELL_GAUSS)
    layout = gpu.plan.layout()
    assert layout["gauss_chunks"] == (int(chunks) if chunks is not None else 1)
    assert layout["brick_cap"] == (int(cap) if cap is not None else layout["brick_cap"])
    seeds = np.random.default_rng(12).uniform(-1.0, 1.0, int(np.prod(gpu.shape)))
    seeds[::5] = 0.0
    pose_grad = mode == "both"
    g_ref, p_ref, _, ref_status = ref.vjp(values, seeds, pose, precision="fp64", pose=pose_grad,
                                          integration="cell_gauss")
    g, p, status = gpu.vjp(values, seeds, pose, pose_grad=pose_grad)
    assert status.tolist() == ref_status.tolist() == [0] * gpu.shape[0]
    assert np.count_nonzero(g_ref) > 0
    np.testing.assert_allclose(g, g_ref, rtol=2e-7, atol=1e-12 * np.abs(g_ref).max())
    if pose_grad:
        np.testing.assert_allclose(p, p_ref, rtol=1e-10, atol=1e-12 * np.abs(p_ref).max())


def test_projector_cell_gauss_is_exact_and_differentiable():
    """Projector(precision="fp64", integration="cell_gauss"): a constant field
    integrates to value x clipped length at any sample count; the midpoint
    rule converges to the cell-Gauss image; autograd gives the adjoint."""
    from dpt import phantoms
    from dpt.torch import Projector

    grid = phantoms.centred_grid((12, 13, 14))
    geometry, _ = phantoms.circular_geometry(grid, 3, (10, 11), sod=60.0, sdd=100.0)
    exact = Projector(grid=grid, geometry=geometry, samples_per_ray=3, precision="fp64", integration="cell_gauss")
    field = torch.tensor(phantoms.chest_phantom(grid, seed=2), device="cuda")
    image = exact(field)
    fine = Projector(grid=grid, geometry=geometry, samples_per_ray=8192, precision="fp64")(field)
    scale = float(image.abs().max())
    assert float((fine - image).abs().max()) <= 1e-5 * scale
    ones = torch.full(grid.shape, 0.5, device="cuda")
    again = Projector(grid=grid, geometry=geometry, samples_per_ray=977, precision="fp64",
                      integration="cell_gauss")(ones)
    torch.testing.assert_close(exact(ones), again, rtol=1e-14, atol=0)
    leaf = field.clone().requires_grad_(True)
    pose = torch.zeros(6, dtype=torch.float64, device="cuda", requires_grad=True)
    out = exact(leaf, pose)
    c = torch.randn(out.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    g_field, g_pose = torch.autograd.grad(out, (leaf, pose), c)
    delta = torch.rand(grid.shape, device="cuda", generator=torch.Generator("cuda").manual_seed(2))
    lhs = float((exact(delta).double() * c).sum())
    rhs = float((delta.double() * g_field.double()).sum())
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), abs(rhs))
    assert torch.isfinite(g_pose).all()
    with pytest.raises(ContractError, match="float64"):
        Projector(grid=grid, geometry=geometry, samples_per_ray=3, integration="cell_gauss")
    with pytest.raises(ContractError, match="midpoint"):
        exact.adjoint_gather(c, pose.detach())


def test_cell_gauss_pose_gradient_matches_central_differences():
    grid, pose, _, gpu = _pair("cone")
    values = _values(grid)
    seeds = np.linspace(-1.0, 1.0, int(np.prod(gpu.shape)))
    _, tangent, status = gpu.vjp(values, seeds, pose, volume=False, components=6)
    assert status.tolist() == [0] * gpu.shape[0]
    direction = np.array([0.2, -0.3, 0.1, 0.04, 0.03, -0.05])

    def loss(p):
        image, _ = gpu.forward(values, p)
        return math.fsum(image.reshape(-1) * seeds)

    eps = 1e-5
    fd = (loss(scenes.compose(pose, eps * direction)) - loss(scenes.compose(pose, -eps * direction))) / (2 * eps)
    analytic = float(tangent.sum(axis=0) @ direction)
    assert abs(fd - analytic) <= 5e-5 * max(1.0, abs(fd))
