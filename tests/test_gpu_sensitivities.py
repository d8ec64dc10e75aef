"""Per-ray pose sensitivities and the pose-fit normal equations (GPU).

dpt_pose_sensitivities exports cotangent x dL/d(delta) per ray for
pose @ exp(delta^) (projection_pose_sensitivities,
python/dpt/projection.py:1021-1084).  The rows are checked ray by ray against
the oracle's per-ray tangent (FP32, FP64 and exact cell-Gauss rules; the
oracle's pose derivatives are pinned by central differences in
tests/test_oracle.py) and their per-view sums against the reduced pose VJP.
The serving solver's first
evaluation (J^T J, J^T r and 1/2 r^2 summed in FP64 by the fit kernel) is
checked against the normal equations built from the oracle's rows and image.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests import scenes

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt import _native, phantoms, serving  # noqa: E402
from dpt.geometry import ConeBeamIntrinsics  # noqa: E402

from .test_gpu_parity import NativeScene, _cone  # noqa: E402


def _sensitivities(gpu: NativeScene, values, seeds, pose):
    field, _ = gpu._field(values)
    poses = gpu._poses(pose)
    dtype = torch.float32 if gpu.precision == "float32" else torch.float64
    seed = torch.as_tensor(np.ascontiguousarray(seeds).reshape(gpu.shape)).to(dtype).cuda()
    out = torch.full((*gpu.shape, 6), 9.0, dtype=torch.float64, device="cuda")
    gpu.plan.pose_sensitivities(field.data_ptr(), poses.data_ptr(), seed.data_ptr(), out.data_ptr(), 0, gpu.stream)
    status = gpu.plan.status(gpu.stream)
    return out.cpu().numpy().reshape(-1, 6), status


def _case(name):
    """(grid, values, pose, geometry kwargs) of the shared parity scenes."""
    if name == "oblique":
        grid, src, end, pose, values = scenes.oblique_scene()
        return grid, values, pose, dict(rays=(src, end))
    if name == "workflow":
        grid, panels, values = scenes.workflow_scene()
        return grid, values, scenes.compose(scenes.IDENTITY12, (0.03, -0.02, 0.01, 0.04, -0.03, 0.02)), \
            dict(panels=panels)
    grid, panels, values = _cone()
    return grid, values, scenes.compose(scenes.IDENTITY12, (0.5, -0.7, 0.3, 0.02, -0.015, 0.03)), dict(panels=panels)


@pytest.mark.parametrize("rule", ["fp32", "fp64", "cell_gauss"])
@pytest.mark.parametrize("name", ["oblique", "workflow", "cone"])
def test_pose_sensitivities_match_the_oracle_ray_by_ray(name, rule):
    print(f"case sensitivities {name} {rule}", flush=True)
    grid, values, pose, geometry = _case(name)
    precision = "float32" if rule == "fp32" else "float64"
    extra = _native.CELL_GAUSS if rule == "cell_gauss" else 0
    ref = O.Scene(grid, 97, **geometry)
    gpu = NativeScene(grid, 97, precision=precision, extra_needs=extra, **geometry)
    fp = "fp32" if rule == "fp32" else "fp64"
    integration = "cell_gauss" if rule == "cell_gauss" else "midpoint"
    image, _ = ref.forward(values, pose, precision=fp, integration=integration)
    seeds = np.random.default_rng(17).uniform(-1.0, 1.0, image.size)
    seeds[::6] = 0.0
    if rule == "fp32":
        seeds = seeds.astype(np.float32)
    _, _, rows, _ = ref.vjp(values, seeds, pose, precision=fp, volume=False, components=6, per_ray=True,
                            integration=integration)
    actual, status = _sensitivities(gpu, values, seeds, pose)
    assert status.tolist() == [0] * gpu.shape[0]
    assert np.all(actual[seeds == 0.0] == 0.0)
    assert np.count_nonzero(rows) > 0
    scale = np.abs(rows).max()
    if rule == "fp32":
        np.testing.assert_allclose(actual, rows, rtol=1e-5, atol=1e-6 * scale)
    else:
        np.testing.assert_allclose(actual, rows, rtol=1e-10, atol=1e-13 * scale)
    # per-view sums are the reduced pose tangent of dpt_vjp (97 samples: one
    # sample segment per ray, so the same FP32 sums)
    _, tangent, status = gpu.vjp(values, seeds, pose, volume=False, components=6)
    per_view = actual.reshape(gpu.shape[0], -1, 6).sum(axis=1)
    np.testing.assert_allclose(per_view, tangent, rtol=1e-9, atol=1e-12 * np.abs(tangent).max())


def test_fit_normal_equations_match_oracle_rows():
    """The fit kernel's first evaluation at the initial pose: 1/2 r^2, J^T r
    and J^T J against the oracle's FP32 image and unit-weight tangent rows."""
    grid = phantoms.centred_grid((36, 36, 36))
    values = phantoms.chest_phantom(grid, seed=12)
    angles = [0.3, 1.9, 4.0]
    det, samples = 32, 64
    pitch = phantoms.coverage_pitch(grid, 150.0, 240.0, angles, (det, det), margin=1.3)
    geometry = ConeBeamIntrinsics(150.0, 240.0, (det, det), (pitch, pitch), centre_mm=grid.centre_mm()).batch(angles)
    panels = [O.Panel(p.source_mm, p.origin_mm, p.u, p.v, p.spacing_mm, p.shape) for p in geometry.views]
    scene = O.Scene(O.Grid(grid.shape, grid.spacing_mm, grid.origin_mm, grid.orientation), samples, panels=panels)
    start = scenes.compose(scenes.IDENTITY12, (0.8, -0.5, 0.3, 0.01, -0.02, 0.015))
    target, _ = scene.forward(values, scenes.compose(scenes.IDENTITY12, (0.0, 0.4, -0.2, 0.0, 0.01, 0.0)))
    image, _ = scene.forward(values, start)
    residual = (image.astype(np.float64) - target.astype(np.float64)).ravel()
    ones = np.ones(image.size, dtype=np.float32)
    _, _, rows, _ = scene.vjp(values, ones, start, volume=False, components=6, per_ray=True)
    jtj = rows.T @ rows
    jtr = rows.T @ residual
    loss = 0.5 * float(residual @ residual)

    registry = serving.VolumeRegistry()
    registry.add("v", values, grid, storage="float32", layout="linear")
    engine = serving.PoseFitEngine(registry, max_requests=2, max_tiles=4096)
    request = serving.FitRequest("a", "v", geometry, target, samples, 1, np.asarray(start))
    with torch.cuda.stream(engine.stream):
        active = [engine._admit(request, 0.0)]
        engine.step(active)
    engine.stream.synchronize()
    state = active[0].state.cpu().numpy()
    engine.close()
    assert state[serving.BEST_LOSS] == pytest.approx(loss, rel=1e-6)
    np.testing.assert_allclose(state[49:55], jtr, rtol=1e-6, atol=1e-6 * np.abs(jtr).max())
    upper = np.asarray([jtj[i, j] for i in range(6) for j in range(i, 6)])
    np.testing.assert_allclose(state[28:49], upper, rtol=1e-6, atol=1e-6 * np.abs(upper).max())
