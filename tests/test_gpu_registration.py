"""Pose recovery on the GPU: recover_parameters over PoseObjective (one
projector forward and pose VJP per evaluation) finds the pose a target was
rendered at; a rotation radius rejects trials outside the chart without
accepting them."""

from __future__ import annotations

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt import phantoms  # noqa: E402
from dpt.registration import RecoveryPolicy, recover_parameters  # noqa: E402
from dpt.torch import PoseObjective, Projector  # noqa: E402


def _scene():
    grid = phantoms.centred_grid((40, 40, 40))
    values = phantoms.chest_phantom(grid, seed=5)
    geometry, _ = phantoms.circular_geometry(grid, 4, (48, 48), sod=250.0, sdd=400.0)
    projector = Projector(grid=grid, geometry=geometry, samples_per_ray=128)
    field = torch.tensor(values, device="cuda")
    return projector, field


def test_pose_recovery_finds_the_rendering_pose():
    projector, field = _scene()
    truth = (1.5, -1.0, 0.8, 0.03, -0.02, 0.025)
    with torch.no_grad():
        target = projector(field, torch.tensor(truth, dtype=torch.float64, device="cuda"))
    scales = (1.0, 1.0, 1.0, 0.01, 0.01, 0.01)
    objective = PoseObjective(projector, field, target, scales=scales)
    result = recover_parameters(objective, (0.0,) * 6, policy=RecoveryPolicy(max_iterations=80,
                                                                               gradient_tolerance=1e-4))
    print(f"case pose recovery: {result.reason} after {result.evaluations} evaluations", flush=True)
    xi = objective.tangent(result.parameters)
    assert result.evaluation.loss < 1e-6 * objective(tuple(0.0 for _ in range(6))).loss
    np.testing.assert_allclose(xi[:3], truth[:3], atol=2e-2)
    np.testing.assert_allclose(xi[3:], truth[3:], atol=2e-4)
    losses = [h.loss for h in result.history]
    assert all(b <= a for a, b in zip(losses, losses[1:]))


def test_rotation_radius_rejects_trials_outside_the_chart():
    projector, field = _scene()
    truth = (0.0, 0.0, 0.0, 0.02, 0.0, 0.0)
    with torch.no_grad():
        target = projector(field, torch.tensor(truth, dtype=torch.float64, device="cuda"))
    objective = PoseObjective(projector, field, target, scales=(1.0, 1.0, 1.0, 0.01, 0.01, 0.01),
                              rotation_radius=0.05)
    result = recover_parameters(objective, (0.0, 0.0, 0.0, 4.5, 0.0, 0.0),
                                policy=RecoveryPolicy(max_iterations=60, gradient_tolerance=1e-4))
    xi = objective.tangent(result.parameters)
    assert math.hypot(*xi[3:]) <= 0.05
    for trial in result.rejected_trials:  # every rejected point lies outside the radius
        rotation = [s * v for s, v in zip(objective.scales[3:], trial.parameters[3:])]
        assert math.hypot(*rotation) > 0.05
    assert xi[3] == pytest.approx(0.02, abs=5e-4)
