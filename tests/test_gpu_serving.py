"""Pose-fit service on the GPU: convergence, batch invariance, consistency
with the Projector's forward and pose VJP, and half-precision storage."""

from __future__ import annotations

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt import phantoms, serving  # noqa: E402
from dpt.geometry import ConeBeamIntrinsics, RigidTransform, compose_pose  # noqa: E402
from dpt.torch import Projector  # noqa: E402


def _scene(n=40, views=3, det=40, seed=0, angles=None):
    grid = phantoms.centred_grid((n, n, n))
    values = phantoms.chest_phantom(grid, seed=seed)
    rng = np.random.default_rng(seed)
    angles = rng.uniform(0, 2 * math.pi, views) if angles is None else angles
    pitch = phantoms.coverage_pitch(grid, 150.0, 240.0, angles, (det, det), margin=1.3)
    geometry = ConeBeamIntrinsics(150.0, 240.0, (det, det), (pitch, pitch), centre_mm=grid.centre_mm()).batch(angles)
    return grid, values, geometry


def _target(grid, values, geometry, xi, samples):
    pose = compose_pose(RigidTransform(), xi).packed()
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)
    with torch.no_grad():
        image = module(torch.tensor(values, device="cuda"), packed_pose=torch.tensor(pose, dtype=torch.float64, device="cuda"))
    return image.cpu().numpy(), pose


def _request(key, volume, geometry, targets, samples, iterations, initial=None):
    return serving.FitRequest(
        key, volume, geometry, targets, samples, iterations,
        RigidTransform().packed() if initial is None else initial,
    )


def test_solve_recovers_the_true_pose_from_exact_data():
    grid, values, geometry = _scene()
    xi = (3.0, -2.0, 1.5, math.radians(2.0), math.radians(-1.5), math.radians(1.0))
    targets, truth = _target(grid, values, geometry, xi, 96)
    registry = serving.VolumeRegistry()
    registry.add("v", values, grid, storage="float32")
    engine = serving.PoseFitEngine(registry, max_requests=4, max_tiles=4096)
    stats = engine.run([_request("a", "v", geometry, targets, 96, 30)], realtime=False)
    (result,) = stats.results
    angle, shift = serving.pose_error(result.pose, truth)
    assert result.evaluations == 30 and result.accepted >= 5
    assert angle < 0.02 and shift < 0.02, (angle, shift, result.loss)
    assert result.loss < 1e-6


def test_results_are_identical_alone_and_in_a_mix_and_with_graphs():
    grid, values, geometry = _scene(seed=1)
    grid2, values2, geometry2 = _scene(n=32, views=5, det=24, seed=2)
    xi = (1.0, 2.0, -1.0, 0.02, 0.01, -0.015)
    t1, _ = _target(grid, values, geometry, xi, 80)
    t2, _ = _target(grid2, values2, geometry2, xi, 48)
    registry = serving.VolumeRegistry()
    registry.add("v1", values, grid, storage="bfloat16")
    registry.add("v2", values2, grid2, storage="float32")
    a = _request("a", "v1", geometry, t1, 80, 12)
    others = [
        _request("b", "v2", geometry2, t2, 48, 7),
        _request("c", "v1", geometry, t1, 80, 12, compose_pose(RigidTransform(), (0.5, 0, 0, 0, 0.01, 0)).packed()),
        _request("d", "v2", geometry2, t2, 48, 20),
    ]
    runs = {}
    for name, batch, graph in (("alone", [a], True), ("mix", [a, *others], True), ("direct", [a, *others], False)):
        engine = serving.PoseFitEngine(registry, max_requests=8, max_tiles=8192, use_graph=graph)
        runs[name] = {r.key: r for r in engine.run(batch, realtime=False).results}
        engine.close()
    for name in ("mix", "direct"):
        assert np.array_equal(runs[name]["a"].pose, runs["alone"]["a"].pose)
        assert runs[name]["a"].loss == runs["alone"]["a"].loss
    for key in ("b", "c", "d"):
        assert np.array_equal(runs["mix"][key].pose, runs["direct"][key].pose)


def test_first_evaluation_matches_projector_loss_and_pose_gradient():
    grid, values, geometry = _scene(seed=3)
    xi = (2.0, 0.0, -1.0, 0.01, 0.02, 0.0)
    targets, _ = _target(grid, values, geometry, xi, 64)
    registry = serving.VolumeRegistry()
    registry.add("v", values, grid, storage="float32")
    engine = serving.PoseFitEngine(registry, max_requests=2, max_tiles=4096)
    request = _request("a", "v", geometry, targets, 64, 1)
    with torch.cuda.stream(engine.stream):
        active = [engine._admit(request, 0.0)]
        engine.step(active)
    engine.stream.synchronize()
    state = active[0].state.cpu().numpy()
    # reference quantities through the Projector at the initial pose
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=64)
    field = torch.tensor(values, device="cuda")
    pose = torch.zeros(6, dtype=torch.float64, device="cuda", requires_grad=True)
    image = module(field, pose)
    residual = image - torch.tensor(targets, device="cuda")
    loss = 0.5 * float((residual.double() ** 2).sum())
    (gradient,) = torch.autograd.grad(image, pose, residual)
    assert state[serving.BEST_LOSS] == pytest.approx(loss, rel=1e-4)
    jtr = state[49:55]
    np.testing.assert_allclose(jtr, gradient.cpu().numpy(), rtol=2e-3, atol=1e-6 * np.abs(jtr).max())


def test_half_storage_solves_converge_like_fp32():
    grid, values, geometry = _scene(seed=4)
    xi = (-2.0, 1.0, 2.0, math.radians(-2.0), math.radians(1.0), math.radians(1.5))
    targets, truth = _target(grid, values, geometry, xi, 96)
    for storage in ("bfloat16", "float16"):
        registry = serving.VolumeRegistry()
        registry.add("v", values, grid, storage=storage)
        engine = serving.PoseFitEngine(registry, max_requests=2, max_tiles=4096)
        (result,) = engine.run([_request("a", "v", geometry, targets, 96, 30)], realtime=False).results
        angle, shift = serving.pose_error(result.pose, truth)
        assert angle < 0.1 and shift < 0.1, (storage, angle, shift)


def test_admission_respects_slots_and_budget():
    grid, values, geometry = _scene(n=24, views=2, det=16, seed=5)
    targets, _ = _target(grid, values, geometry, (0.5, 0, 0, 0, 0, 0.01), 32)
    registry = serving.VolumeRegistry()
    registry.add("v", values, grid, storage="bfloat16")
    requests = [_request(f"r{i}", "v", geometry, targets, 32, 3 + i) for i in range(6)]
    engine = serving.PoseFitEngine(registry, max_requests=2, max_tiles=4096, hbm_budget=2 * requests[0].nbytes)
    stats = engine.run(requests, realtime=False)
    assert len(stats.results) == 6 and stats.peak_active <= 2
    assert stats.max_bytes <= 2 * requests[0].nbytes
    assert sorted(r.key for r in stats.results) == sorted(q.key for q in requests)


@pytest.mark.parametrize("storage", ["float32", "bfloat16", "float16"])
def test_cell_layout_volumes_solve_identically_to_linear(storage):
    """The cell layout moves where corner values are read from, not what is
    computed: every solve is bit-identical to the linear-layout one."""
    grid, values, geometry = _scene(n=36, views=3, det=32, seed=6)
    grid2, values2, geometry2 = _scene(n=28, views=2, det=24, seed=7)
    xi = (1.5, -1.0, 0.5, 0.015, -0.01, 0.02)
    t1, _ = _target(grid, values, geometry, xi, 72)
    t2, _ = _target(grid2, values2, geometry2, xi, 40)
    results = {}
    for layout in ("linear", "cells"):
        registry = serving.VolumeRegistry()
        registry.add("v1", values, grid, storage=storage, layout=layout)
        registry.add("v2", values2, grid2, storage=storage, layout=layout)
        engine = serving.PoseFitEngine(registry, max_requests=4, max_tiles=8192)
        batch = [_request("a", "v1", geometry, t1, 72, 10), _request("b", "v2", geometry2, t2, 40, 14)]
        results[layout] = {r.key: r for r in engine.run(batch, realtime=False).results}
        engine.close()
    for key in ("a", "b"):
        assert np.array_equal(results["cells"][key].pose, results["linear"][key].pose)
        assert results["cells"][key].loss == results["linear"][key].loss
