"""Projector workflow on the GPU: autograd against the oracle, adjoint and
finite-difference checks, typed device errors with recovery, ``out=``,
saved-tensor mutation, streams and dispatcher tracing.

Mirrors the reference's workflow tests (tests/python/test_projector_workflow.py:
61-138, tests/python/test_torch_projection_registration.py:198-227,
benchmarks/fullscan_guards.py) with this build's oracle as the checker.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt import ContractError, NumericalError, TrialDomainError, phantoms  # noqa: E402
from dpt.geometry import RigidTransform  # noqa: E402
from dpt.torch import Projector  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _scene(n=24, views=3, det=24, samples=96, seed=0):
    grid = phantoms.centred_grid((n, n, n))
    values = phantoms.chest_phantom(grid, seed=seed)
    geometry, _ = phantoms.circular_geometry(grid, views, (det, det), sod=200.0, sdd=300.0)
    return grid, values, geometry, samples


def _oracle(grid, geometry, samples):
    return O.Scene(
        O.Grid(grid.shape, grid.spacing_mm, grid.origin_mm, grid.orientation),
        samples,
        panels=[O.Panel(p.source_mm, p.origin_mm, p.u, p.v, p.spacing_mm, p.shape) for p in geometry.views],
    )


def test_autograd_through_the_projector_matches_the_oracle():
    grid, values, geometry, samples = _scene()
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)
    field = torch.tensor(values, device="cuda", requires_grad=True)
    pose = torch.zeros(6, dtype=torch.float64, device="cuda", requires_grad=True)
    image = module(field, pose)
    seed = torch.cos(torch.arange(image.numel(), device="cuda", dtype=torch.float32) * 0.01).reshape(image.shape)
    loss = (image * seed).sum()
    g_field, g_pose = torch.autograd.grad(loss, (field, pose))
    ref = _oracle(grid, geometry, samples)
    expected, _ = ref.forward(values)
    assert O.relative_l2(image.detach().cpu().numpy(), expected) <= 1e-6
    g_ref, _, _, _ = ref.vjp(values, seed.cpu().numpy())
    assert O.relative_l2(g_field.cpu().numpy().ravel(), g_ref) <= 1e-5
    _, tangent, _, _ = ref.vjp(values, seed.cpu().numpy(), volume=False, components=6)
    assert O.relative_l2(g_pose.cpu().numpy(), tangent.sum(axis=0)) <= 1e-5


@pytest.mark.parametrize("precision,tolerance", [("fp64", 1e-6), ("fp32", 2e-5)])
def test_volume_adjoint_identity(precision, tolerance):
    """<A d, c> = <d, A^T c> for random d and c (adjoint quality gate); the
    volume gradient is returned in the field's FP32 storage in both modes."""
    grid, values, geometry, samples = _scene(n=20, views=4, det=20, samples=80)
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples, precision=precision)
    gen = torch.Generator("cuda").manual_seed(5)
    delta = torch.rand(grid.shape, device="cuda", generator=gen)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    field = delta.clone().requires_grad_(True)
    image = module(field)
    c = torch.randn(image.shape, device="cuda", generator=gen, dtype=dtype)
    (adjoint,) = torch.autograd.grad(image, field, c)
    lhs = float((image.double() * c.double()).sum())
    rhs = float((delta.double() * adjoint.double()).sum())
    assert abs(lhs - rhs) <= tolerance * max(abs(lhs), abs(rhs))


def test_pose_gradient_matches_central_differences_in_fp64():
    grid, values, geometry, samples = _scene(n=20, views=2, det=16, samples=64, seed=3)
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples, precision="fp64")
    field = torch.tensor(values, device="cuda")
    xi0 = torch.tensor([0.7, -0.4, 0.3, 0.02, -0.015, 0.01], dtype=torch.float64, device="cuda")
    seed = torch.linspace(-1.0, 1.0, module.image_shape[0] * 16 * 16, dtype=torch.float64, device="cuda")
    seed = seed.reshape(module.image_shape)
    pose = xi0.clone().requires_grad_(True)
    (grad,) = torch.autograd.grad((module(field, pose) * seed).sum(), pose)
    scale = float(grad.abs().max())
    for axis, h in enumerate((1e-4, 1e-4, 1e-4, 1e-6, 1e-6, 1e-6)):
        e = torch.zeros(6, dtype=torch.float64, device="cuda")
        e[axis] = h
        with torch.no_grad():
            plus = float((module(field, xi0 + e) * seed).sum())
            minus = float((module(field, xi0 - e) * seed).sum())
        fd = (plus - minus) / (2 * h)
        assert math.isclose(float(grad[axis]), fd, rel_tol=2e-3, abs_tol=1e-5 * scale), (axis, float(grad[axis]), fd)


def test_device_errors_are_typed_and_the_next_call_recovers():
    grid, values, geometry, samples = _scene(n=16, views=2, det=16, samples=32)
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)
    good = torch.tensor(values, device="cuda")
    bad = good.clone()
    bad[3, 4, 5] = float("nan")
    with pytest.raises(TrialDomainError, match="finite and non-negative"):
        module(bad)
    negative = good.clone()
    negative[1, 1, 1] = -1.0
    with pytest.raises(TrialDomainError):
        module(negative)
    image = module(good)  # status words were consumed: this call succeeds
    assert torch.isfinite(image).all()
    scaled = torch.tensor(RigidTransform().packed(), dtype=torch.float64, device="cuda")
    scaled[0] = 1.01  # not a rotation
    with pytest.raises(ContractError, match="rigid"):
        module(good, packed_pose=scaled)
    huge = good.clone()
    huge[8, 8, 8] = 3.0e38
    with pytest.raises(NumericalError):
        module(huge)
    assert torch.isfinite(module(good)).all()


def test_out_argument_writes_in_place_and_rejects_bad_buffers():
    grid, values, geometry, samples = _scene(n=16, views=2, det=16, samples=32)
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)
    field = torch.tensor(values, device="cuda")
    out = torch.empty(module.image_shape, device="cuda")
    with torch.no_grad():
        result = module(field, out=out)
    assert result is out
    torch.testing.assert_close(out, module(field), rtol=0, atol=0)
    with pytest.raises(ContractError):
        module(field, out=torch.empty(module.image_shape, device="cuda", dtype=torch.float64))
    with pytest.raises(ContractError):
        module(field.requires_grad_(True), out=out)


def test_saved_field_mutation_is_detected_before_backward():
    grid, values, geometry, samples = _scene(n=16, views=2, det=16, samples=32)
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)
    field = torch.tensor(values, device="cuda", requires_grad=True)
    leaf = field * 1.0
    image = module(leaf)
    with torch.no_grad():
        leaf.mul_(2.0)
    with pytest.raises(RuntimeError, match="modified by an inplace operation"):
        image.sum().backward()


def test_side_stream_gives_the_same_result():
    grid, values, geometry, samples = _scene(n=16, views=3, det=16, samples=48)
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)
    field = torch.tensor(values, device="cuda")
    expected = module(field)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        got = module(field)
    torch.cuda.current_stream().wait_stream(side)
    torch.testing.assert_close(got, expected, rtol=0, atol=0)


def test_torch_compile_fullgraph_traces_the_custom_ops():
    grid, values, geometry, samples = _scene(n=16, views=2, det=16, samples=32)
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)
    field = torch.tensor(values, device="cuda", requires_grad=True)

    def loss(f):
        return (module(f) ** 2).sum()

    compiled = torch.compile(loss, fullgraph=True, backend="aot_eager")
    eager_value = loss(field)
    (eager_grad,) = torch.autograd.grad(eager_value, field)
    value = compiled(field)
    (grad,) = torch.autograd.grad(value, field)
    torch.testing.assert_close(value, eager_value, rtol=1e-6, atol=0)
    torch.testing.assert_close(grad, eager_grad, rtol=1e-6, atol=1e-12)


def test_api_project_matches_the_module():
    import dpt

    grid, values, geometry, samples = _scene(n=16, views=2, det=16, samples=40)
    volume = dpt.ImagingVolume(np.asarray(values), grid)
    image = dpt.project(volume, geometry, device="cuda:0", samples_per_ray=samples)
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)
    torch.testing.assert_close(image, module(torch.tensor(values, device="cuda")), rtol=0, atol=0)


def _chart_cases():
    rng = np.random.default_rng(61)
    tangents = [np.zeros(6), np.full(6, 1e-9), np.array([3.0, -2.0, 1.0, 0.3, -0.2, 0.1]),
                np.array([120.0, -40.0, 7.0, 2.9, -1.3, 0.4]), np.array([0.0, 0.0, 0.0, 0.0, 0.0, math.pi]),
                np.array([5.0, 1.0, -2.0, 7.0, -9.0, 11.0])]
    tangents += [np.concatenate([rng.uniform(-50, 50, 3), rng.uniform(-4, 4, 3)]) for _ in range(10)]
    return np.stack(tangents)


def test_device_pose_chart_matches_matrix_exp_and_its_gradients():
    """dpt::compose_pose (one HIP launch) against the matrix_exp chart of
    python/dpt/torch/geometry.py:19-63 on host tensors: values and the
    gradients of anchor and tangent, for zero, tiny, large-angle (pi and
    beyond) and random twists, with a broadcast anchor."""
    from dpt.torch import compose_pose

    tangents = _chart_cases()
    base = compose_pose(torch.tensor(RigidTransform().packed(), dtype=torch.float64),
                        torch.tensor([10.0, -20.0, 5.0, 0.4, -0.7, 1.1], dtype=torch.float64))
    anchors = compose_pose(base.expand(len(tangents), 12),
                           torch.tensor(np.random.default_rng(62).uniform(-1, 1, (len(tangents), 6))))
    cot = torch.tensor(np.random.default_rng(63).normal(size=(len(tangents), 12)))
    for anchor_host in (anchors, base):  # per-pose anchors, one broadcast anchor
        a_host = anchor_host.clone().requires_grad_(True)
        x_host = torch.tensor(tangents, requires_grad=True)
        out_host = compose_pose(a_host, x_host)
        ga_host, gx_host = torch.autograd.grad(out_host, (a_host, x_host), cot)
        a_dev = anchor_host.cuda().requires_grad_(True)
        x_dev = torch.tensor(tangents, device="cuda", requires_grad=True)
        out_dev = compose_pose(a_dev, x_dev)
        ga_dev, gx_dev = torch.autograd.grad(out_dev, (a_dev, x_dev), cot.cuda())
        assert out_dev.shape == out_host.shape
        torch.testing.assert_close(out_dev.cpu(), out_host.detach(), rtol=1e-13, atol=1e-12)
        torch.testing.assert_close(ga_dev.cpu(), ga_host, rtol=1e-12, atol=1e-11)
        torch.testing.assert_close(gx_dev.cpu(), gx_host, rtol=1e-11, atol=1e-10)


def test_device_pose_chart_gradcheck():
    from dpt.torch import compose_pose

    anchor = torch.tensor(RigidTransform().packed(), dtype=torch.float64, device="cuda")
    anchor = compose_pose(anchor, torch.tensor([1.0, 2.0, 3.0, 0.2, -0.1, 0.3], dtype=torch.float64,
                                                device="cuda")).detach().requires_grad_(True)
    tangent = torch.tensor([[0.5, -0.3, 0.2, 0.7, -0.4, 0.25], [0.0, 0.0, 0.0, 0.0, 0.0, 0.0]],
                           dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(compose_pose, (anchor, tangent), eps=1e-6, atol=1e-7, rtol=1e-6)
