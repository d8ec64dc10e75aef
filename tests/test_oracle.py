"""Pin the CPU oracle before it is trusted as the checker (no GPU).

Every check below compares the oracle against something it does not share code
with: closed forms transcribed from the reference's own tests, independent
NumPy restatements (explicit FP32 matrix, exact piecewise-Gauss integral),
adjoint identities, central finite differences and analytic ellipsoid chords.
"""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from tests import scenes

GOLDEN = json.loads((Path(__file__).parent / "golden" / "closed_forms.json").read_text())


def _grid(spec):
    return O.Grid(tuple(spec["shape"]), tuple(spec["spacing_mm"]), tuple(spec.get("origin_mm", (0.0, 0.0, 0.0))))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("case", GOLDEN["projector"], ids=lambda c: c["name"])
def test_oracle_matches_reference_closed_form(case, precision):
    grid = _grid(case["grid"])
    rays = (np.asarray(case["rays"]["sources"]), np.asarray(case["rays"]["endpoints"]))
    scene = O.Scene(grid, case["samples"], rays=rays)
    image, status = scene.forward(case["field"], precision=precision)
    assert status.tolist() == [0]
    np.testing.assert_allclose(image.ravel(), case["forward"], rtol=case["rtol"])
    volume, tangent, _, status = scene.vjp(case["field"], case["seed"], precision=precision, components=6)
    assert status.tolist() == [0]
    np.testing.assert_allclose(volume, case["volume_vjp"], rtol=case["rtol"])
    np.testing.assert_allclose(tangent.ravel(), case["pose_tangent"], rtol=case["rtol"], atol=case["atol_pose"])


@pytest.mark.parametrize("case", GOLDEN["integrals"], ids=lambda c: c["name"])
def test_exact_integral_restatement_matches_closed_forms(case):
    grid = _grid(case["grid"])
    source, endpoint = case["segment"]
    value = O.integrate_segment(case["field"], grid, None, source, endpoint)
    assert value == pytest.approx(case["exact"], rel=1e-12, abs=1e-15)


@pytest.mark.parametrize("case", GOLDEN["integrals"], ids=lambda c: c["name"])
def test_fp64_midpoint_oracle_converges_to_exact_integral(case):
    grid = _grid(case["grid"])
    source, endpoint = case["segment"]
    rays = (np.asarray([source]), np.asarray([endpoint]))
    errors = []
    for samples in (64, 256, 1024):
        image, _ = O.Scene(grid, samples, rays=rays).forward(case["field"], precision="fp64")
        errors.append(abs(float(image.ravel()[0]) - case["exact"]))
    assert errors[-1] <= 1e-5 * max(1.0, abs(case["exact"]))
    assert errors[-1] <= errors[0] + 1e-15


@pytest.mark.parametrize("case", GOLDEN["samples"], ids=lambda c: c["name"])
def test_sample_restatement_matches_closed_forms(case):
    grid = O.Grid(tuple(case["grid"]["shape"]), tuple(case["grid"]["spacing_mm"]))
    if "field" in case:
        values = case["field"]
    else:
        nz, ny, nx = grid.shape
        values = [1.0 + x + 2 * y + 3 * z for z in range(nz) for y in range(ny) for x in range(nx)]
    assert O.sample_reference(values, grid, case["point"]) == pytest.approx(case["value"], abs=1e-12)


@pytest.mark.parametrize("case", GOLDEN["ellipsoids"], ids=lambda c: c["name"])
def test_ellipsoid_chord_restatement(case):
    source, target = case["segment"]
    assert O.ellipsoid_chord(case["axes"], case["mu"], case["centre"], source, target) == pytest.approx(case["depth"])


@pytest.mark.parametrize("samples", [17, 97, 1024])
def test_fp32_oracle_matches_independent_explicit_matrix(samples):
    """tests/python/warp_cpu/test_projection_fp32.py:201-213 restated."""
    grid, sources, endpoints, pose, values = scenes.oblique_scene()
    matrix = O.explicit_fp32_matrix(grid, list(zip(sources, endpoints)), pose, samples)
    scene = O.Scene(grid, samples, rays=(sources, endpoints))
    image, status = scene.forward(values, pose)
    assert status.tolist() == [0]
    actual = image.ravel().astype(np.float64)
    np.testing.assert_allclose(actual, matrix @ values, rtol=2e-5, atol=3e-7)
    seeds = np.asarray([0.7, -0.3, 0.2, 0.6, 1.0], dtype=np.float32)
    volume, _, _, _ = scene.vjp(values, seeds, pose, pose=False)
    np.testing.assert_allclose(volume, matrix.T @ seeds, rtol=2e-5, atol=3e-7)
    left, right = float(actual @ seeds), float(values.astype(np.float64) @ volume)
    assert abs(left - right) / max(abs(left), abs(right)) < 1e-5


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_oracle_matches_independent_47_midpoint_matrix(precision):
    """tests/python/warp_cpu/test_projector_workflow.py:61-84 restated."""
    grid, panels, field = scenes.workflow_scene()
    columns = []
    for voxel in range(grid.voxels):
        basis = np.zeros(grid.voxels)
        basis[voxel] = 1.0
        columns.append(
            [
                O.integrate_segment(basis, grid, None, p.source, p.pixel(r, c), midpoint_samples=47)
                for p in panels
                for r in range(p.shape[0])
                for c in range(p.shape[1])
            ]
        )
    operator = np.asarray(columns).T
    scene = O.Scene(grid, 47, panels=panels)
    image, status = scene.forward(field, precision=precision)
    assert status.tolist() == [0, 0]
    tolerance = 5e-6 if precision == "fp32" else 2e-11
    np.testing.assert_allclose(image.ravel(), operator @ field.astype(np.float64), rtol=tolerance, atol=tolerance)
    seed = np.linspace(-0.5, 0.8, image.size)
    volume, _, _, _ = scene.vjp(field, seed, precision=precision, pose=False)
    np.testing.assert_allclose(volume, operator.T @ seed, rtol=5e-6, atol=5e-7)


def _loss(scene, values, pose, seeds, precision="fp64"):
    image, _ = scene.forward(values, pose, precision=precision)
    return float(image.ravel().astype(np.float64) @ np.asarray(seeds, dtype=np.float64))


def test_fp32_pose_tangent_matches_fp64_and_finite_differences():
    """tests/python/warp_cpu/test_projection_fp32.py:216-245 restated."""
    grid, sources, endpoints, pose, values = scenes.oblique_scene()
    scene = O.Scene(grid, 97, rays=(sources, endpoints))
    seeds = np.asarray([0.7, -0.3, 0.2, 0.6, 0.0], dtype=np.float32)
    _, g32, _, _ = scene.vjp(values, seeds, pose, volume=False, components=6)
    _, g64, _, _ = scene.vjp(values, seeds, pose, precision="fp64", volume=False, components=6)
    np.testing.assert_allclose(g32.ravel(), g64.ravel(), rtol=2e-3, atol=1e-5)
    for direction in (np.array([0.4, -0.2, 0.3, 0.01, -0.02, 0.015]), np.array([0.02, 0.03, -0.04, 0.15, -0.11, 0.07])):
        expected = float(g64.ravel() @ direction)
        errors = []
        for eps in (3e-3, 1e-3, 3e-4):
            plus = _loss(scene, values, scenes.compose(pose, eps * direction), seeds)
            minus = _loss(scene, values, scenes.compose(pose, -eps * direction), seeds)
            errors.append(abs((plus - minus) / (2 * eps) - expected) / max(abs(expected), 1e-12))
        assert min(errors) < 1e-3, (expected, errors)


def test_fp64_packed_pose_gradient_matches_finite_differences_per_entry():
    grid, panels, field = scenes.workflow_scene()
    scene = O.Scene(grid, 47, panels=panels)
    pose = scenes.compose(scenes.IDENTITY12, (0.03, -0.02, 0.01, 0.04, -0.03, 0.02))
    seeds = np.linspace(-0.5, 0.8, 12)
    _, g, _, _ = scene.vjp(field, seeds, pose, precision="fp64", volume=False, components=12)
    # translation entries are unconstrained coordinates: central differences
    for k in (9, 10, 11):
        eps = 1e-5
        plus, minus = list(pose), list(pose)
        plus[k] += eps
        minus[k] -= eps
        fd = (_loss(scene, field, plus, seeds) - _loss(scene, field, minus, seeds)) / (2 * eps)
        assert fd == pytest.approx(float(g[:, k].sum()), rel=1e-5, abs=1e-9)


def test_per_ray_sum_and_packed_to_tangent_chain():
    """tests/python/warp_cpu/test_projection_fp32.py:290-307 restated."""
    grid, sources, endpoints, pose, values = scenes.oblique_scene()
    scene = O.Scene(grid, 97, rays=(sources, endpoints))
    seeds = np.asarray([0.7, -0.3, 0.2, 0.6, 0.0], dtype=np.float32)
    _, tangent, per_ray, _ = scene.vjp(values, seeds, pose, volume=False, components=6, per_ray=True)
    np.testing.assert_allclose(per_ray.sum(axis=0), tangent.ravel(), rtol=2e-12, atol=2e-12)
    _, packed, _, _ = scene.vjp(values, seeds, pose, volume=False, components=12)
    rotation = np.asarray(pose[:9]).reshape(3, 3)
    np.testing.assert_allclose(rotation.T @ packed.ravel()[9:], tangent.ravel()[:3], rtol=2e-12, atol=2e-12)


def test_large_origin_invariance():
    """tests/python/warp_cpu/test_projection_fp32.py:260-287 restated."""
    grid, sources, endpoints, _, values = scenes.oblique_scene()
    reference, _ = O.Scene(grid, 97, rays=(sources, endpoints)).forward(values)
    shift = np.array([1e8, -2e8, 3e8])
    moved = O.Grid(grid.shape, grid.spacing, tuple(np.asarray(grid.origin) + shift), grid.orientation)
    image, _ = O.Scene(moved, 97, rays=(sources + shift, endpoints + shift)).forward(values)
    np.testing.assert_allclose(image, reference, rtol=2e-6, atol=2e-7)


def test_constant_field_gives_exact_length_and_zero_slope():
    """tests/python/warp_cpu/test_projection_fp32_slopes.py:58-70 (constant field)."""
    grid = O.Grid((4, 5, 6), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    values = np.full(grid.voxels, 0.03, dtype=np.float32)
    sources = np.array([[-10.0, 2.0, 1.5]])
    endpoints = np.array([[20.0, 2.0, 1.5]])
    scene = O.Scene(grid, 64, rays=(sources, endpoints))
    image, _ = scene.forward(values, precision="fp64")
    # chord through the support box x in [-0.5, 5.5]: 6 mm
    assert float(image.ravel()[0]) == pytest.approx(float(np.float32(0.03)) * 6.0, rel=1e-12)


def test_analytic_sphere_chords_converge_under_refinement():
    """tests/python/warp_cpu/test_analytic_phantom_projection.py:18-64 (idea)."""
    radius, mu = 10.0, 0.02
    source, target = (-40.0, 0.7, -1.3), (40.0, -0.4, 2.1)
    exact = O.ellipsoid_chord((radius,) * 3, mu, (0.0, 0.0, 0.0), source, target)
    errors = []
    for n, h in ((16, 1.6), (32, 0.8), (64, 0.4)):
        grid = O.Grid((n, n, n), (h, h, h), (-(n - 1) * h / 2,) * 3)
        nx = np.arange(n) * h - (n - 1) * h / 2
        z, y, x = np.meshgrid(nx, nx, nx, indexing="ij")
        values = ((x * x + y * y + z * z) <= radius * radius).astype(np.float32) * mu
        image, _ = O.Scene(grid, 2048, rays=(np.array([source]), np.array([target]))).forward(values, precision="fp64")
        errors.append(abs(float(image.ravel()[0]) - exact))
    assert errors[2] < errors[0]
    assert errors[2] / exact < 0.05


def test_planted_mistakes_are_detected_by_the_pins():
    """A dropped clip-face term or a transposed pose fails the FD / matrix pins."""
    grid, sources, endpoints, pose, values = scenes.oblique_scene()
    scene = O.Scene(grid, 97, rays=(sources, endpoints))
    seeds = np.asarray([0.7, -0.3, 0.2, 0.6, 0.0], dtype=np.float32)
    _, g64, _, _ = scene.vjp(values, seeds, pose, precision="fp64", volume=False, components=6)
    # (1) transposed rotation in the forward model changes the loss far beyond FD noise
    transposed = tuple(np.asarray(pose[:9]).reshape(3, 3).T.reshape(-1)) + tuple(pose[9:])
    assert abs(_loss(scene, values, transposed, seeds) - _loss(scene, values, pose, seeds)) > 1e-4
    # (2) a gradient with its rotation block sign flipped fails the FD check
    wrong = g64.ravel().copy()
    wrong[3:] *= -1
    direction = np.array([0.02, 0.03, -0.04, 0.15, -0.11, 0.07])
    eps = 1e-3
    fd = (
        _loss(scene, values, scenes.compose(pose, eps * direction), seeds)
        - _loss(scene, values, scenes.compose(pose, -eps * direction), seeds)
    ) / (2 * eps)
    assert abs(fd - wrong @ direction) / abs(fd) > 1e-2
    assert abs(fd - g64.ravel() @ direction) / abs(fd) < 1e-3


def test_oracle_flags_invalid_rays_and_poses():
    grid = O.Grid((2, 2, 2), (1.0, 1.0, 1.0))
    scene = O.Scene(grid, 8, rays=(np.array([[-5.0, 0.5, 0.5]]), np.array([[5.0, 0.5, 0.5]])))
    bad_pose = list(scenes.IDENTITY12)
    bad_pose[9] = math.nan
    _, status = scene.forward(np.ones(8, dtype=np.float32), bad_pose)
    assert status.tolist() == [2]


# ---------------------------------------------------------------------------
# Deterministic adjoint gather (kernels/projection_adjoint.py:73-170)
# ---------------------------------------------------------------------------


def _gather_case(case):
    if case == "workflow":
        grid, panels, _ = scenes.workflow_scene(3)
        return grid, dict(panels=panels), scenes.compose(scenes.IDENTITY12, (0.03, -0.02, 0.01, 0.04, -0.03, 0.02))
    return scenes.gather_scene(case)


GATHER_ORACLE_CASES = [*scenes.GATHER_CASES, "workflow"]


@pytest.mark.parametrize("case", GATHER_ORACLE_CASES)
def test_gather_serial_order_equals_the_voxel_driven_restatement(case):
    """The serial (view, pixel, sample, corner) scatter hands each voxel the
    addends in the reference's per-voxel order, so the sums are identical."""
    grid, geometry, pose = _gather_case(case)
    scene = O.Scene(grid, 31, **geometry)
    seeds = np.random.default_rng(584291).uniform(-1.0, 1.0, scene.views * scene.rows * scene.cols)
    fast, status = scene.gather(seeds, pose)
    assert status.tolist() == [0] * scene.views
    slow = O.gather_voxel_driven(scene, seeds, pose)
    np.testing.assert_array_equal(fast.view(np.int64), slow.view(np.int64))
    assert np.count_nonzero(fast) > 0


def _forward_matrix(scene, grid, pose):
    columns = []
    for voxel in range(grid.voxels):
        basis = np.zeros(grid.voxels, dtype=np.float32)
        basis[voxel] = 1.0
        image, _ = scene.forward(basis, pose, precision="fp64")
        columns.append(image.reshape(-1))
    return np.stack(columns, axis=1)  # (pixels, voxels)


@pytest.mark.parametrize("case", GATHER_ORACLE_CASES)
def test_gather_is_the_transpose_of_the_fp64_forward_and_repeatable(case):
    """tests/python/warp_cpu/test_projection_adjoint_gather.py:97-155: the
    gather equals sum_p A[p,v] c[p] over the explicit FP64 forward matrix
    (2e-12), satisfies <A mu, c> = <mu, A^T c>, repeats bit for bit and agrees
    with the FP64 scatter VJP (whose increments are FP32-rounded)."""
    grid, geometry, pose = _gather_case(case)
    scene = O.Scene(grid, 31, **geometry)
    matrix = _forward_matrix(scene, grid, pose)
    rng = np.random.default_rng(584291)
    for _ in range(3):
        values = rng.uniform(-1.0, 1.0, grid.voxels).astype(np.float32)
        seeds = rng.uniform(-1.0, 1.0, matrix.shape[0])
        actual, _ = scene.gather(seeds, pose)
        expected = np.array([math.fsum(matrix[:, v] * seeds) for v in range(grid.voxels)])
        np.testing.assert_allclose(actual, expected, rtol=2e-12, atol=2e-12)
        image, _ = scene.forward(values, pose, precision="fp64")
        left = math.fsum(image.reshape(-1) * seeds)
        right = math.fsum(values.astype(np.float64) * actual)
        assert abs(left - right) < 2e-12 * max(1.0, abs(left), abs(right))
        again, _ = scene.gather(seeds, pose)
        assert again.tobytes() == actual.tobytes()
        scatter, _, _, _ = scene.vjp(values, seeds, pose, precision="fp64", pose=False)
        np.testing.assert_allclose(scatter, actual, rtol=1e-6, atol=1e-7 * np.abs(actual).max())


def test_gather_accumulates_and_rounds_once_to_float32():
    """tests/python/warp_cpu/test_projection_adjoint_gather.py:158-195."""
    grid, geometry, pose = scenes.gather_scene("clamped")
    scene = O.Scene(grid, 31, **geometry)
    ones = np.ones(scene.views * scene.rows * scene.cols)
    first, _ = scene.gather(ones, pose)
    second, _ = scene.gather(ones, pose, previous=first)
    assert second.tobytes() == (2.0 * first).tobytes()
    single, _ = scene.gather(ones, pose, precision="float32")
    assert single.tobytes() == first.astype(np.float32).tobytes()
    again, _ = scene.gather(ones, pose, previous=single, precision="float32")
    assert again.tobytes() == (first + single.astype(np.float64)).astype(np.float32).tobytes()


def test_gather_flags_overflow_and_invalid_rays():
    """tests/python/warp_cpu/test_projection_adjoint_gather.py:219-235 (1e308
    cotangent accumulated onto 1e308) and an invalid pose translation."""
    grid = O.Grid((1, 1, 1), (1.0, 1.0, 1.0))
    scene = O.Scene(grid, 3, rays=(np.array([[0.0, 0.0, -2.0]]), np.array([[0.0, 0.0, 2.0]])))
    out, status = scene.gather([1e308], previous=np.array([1e308]))
    assert status.tolist() == [2] and not np.isfinite(out).all()
    plain, status = scene.gather([1.0])
    assert status.tolist() == [0] and plain[0] > 0.0
    bad = list(scenes.IDENTITY12)
    bad[9] = math.nan
    _, status = scene.gather([1.0], bad)
    assert status.tolist() == [2]


# ---------------------------------------------------------------------------
# Exact cell-Gauss integration (kernels/projection.py:522-746)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("case", GOLDEN["integrals"], ids=lambda c: c["name"])
def test_cell_gauss_oracle_is_exact_on_closed_forms(case):
    """Two Gauss nodes per cell piece integrate the piecewise-cubic trilinear
    line integral exactly: the closed forms hold to rounding, at any sample
    count (the rule does not use it)."""
    grid = _grid(case["grid"])
    source, endpoint = case["segment"]
    rays = (np.asarray([source]), np.asarray([endpoint]))
    for samples in (1, 7):
        image, status = O.Scene(grid, samples, rays=rays).forward(case["field"], precision="fp64",
                                                                  integration="cell_gauss")
        assert status.tolist() == [0]
        assert float(image.ravel()[0]) == pytest.approx(case["exact"], rel=1e-12, abs=1e-14)


def _gauss_scenes():
    grid, sources, endpoints, pose, values = scenes.oblique_scene()
    yield "oblique", grid, dict(rays=(sources, endpoints)), pose, values
    cgrid = O.Grid((6, 7, 8), (1.0, 0.9, 1.1), (-3.0, -3.5, -2.5))
    panels = scenes.cone_panels(cgrid, 3, 5, 6, sod=40.0, sdd=70.0, offset=0.3)
    yield "cone", cgrid, dict(panels=panels), scenes.compose(scenes.IDENTITY12, (0.3, -0.2, 0.1, 0.05, -0.04, 0.02)), \
        scenes.smooth_field(cgrid, 5)


@pytest.mark.parametrize("name", ["oblique", "cone"])
def test_cell_gauss_oracle_matches_the_independent_exact_integral(name):
    """Against integrate_segment (validation/projection.py:204-251, pure
    Python: breaks at every centre plane, Gauss per piece) ray by ray."""
    for scene_name, grid, geometry, pose, values in _gauss_scenes():
        if scene_name != name:
            continue
        scene = O.Scene(grid, 5, **geometry)
        image, status = scene.forward(values, pose, precision="fp64", integration="cell_gauss")
        assert status.tolist() == [0] * scene.views
        flat = image.reshape(-1)
        if "rays" in geometry:
            segments = list(zip(*geometry["rays"]))
        else:
            segments = [(panel.source, panel.pixel(r, c)) for panel in geometry["panels"]
                        for r in range(panel.shape[0]) for c in range(panel.shape[1])]
        scale = np.abs(flat).max()
        for i, (src, end) in enumerate(segments):
            exact = O.integrate_segment(values, grid, pose, src, end)
            assert abs(flat[i] - exact) <= 1e-12 * scale, (i, flat[i], exact)
        midpoint, _ = O.Scene(grid, 4096, **geometry).forward(values, pose, precision="fp64")
        assert np.abs(midpoint.reshape(-1) - flat).max() <= 1e-5 * scale


@pytest.mark.parametrize("name", ["oblique", "cone"])
def test_cell_gauss_vjp_is_the_adjoint_and_the_pose_derivative(name):
    """<A mu, c> = <mu, A^T c> (the volume gradient sums FP32-rounded
    increments: 1e-6), and the pose tangent matches central differences of
    the exact forward (5e-5; the integral is smooth in the pose)."""
    for scene_name, grid, geometry, pose, values in _gauss_scenes():
        if scene_name != name:
            continue
        scene = O.Scene(grid, 5, **geometry)
        rng = np.random.default_rng(12)
        mu = rng.uniform(0.0, 1.0, grid.voxels).astype(np.float32)
        seeds = rng.uniform(-1.0, 1.0, scene.views * scene.rows * scene.cols)
        image, _ = scene.forward(mu, pose, precision="fp64", integration="cell_gauss")
        g, tangent, _, status = scene.vjp(values, seeds, pose, precision="fp64", components=6,
                                          integration="cell_gauss")
        assert status.tolist() == [0] * scene.views
        g_mu, _, _, _ = scene.vjp(mu, seeds, pose, precision="fp64", pose=False, integration="cell_gauss")
        left = math.fsum(image.reshape(-1) * seeds)
        right = math.fsum(mu.astype(np.float64) * g_mu)
        assert abs(left - right) <= 1e-6 * max(abs(left), abs(right))

        def loss(p):
            out, _ = scene.forward(values, p, precision="fp64", integration="cell_gauss")
            return math.fsum(out.reshape(-1) * seeds)

        direction = np.array([0.2, -0.3, 0.1, 0.04, 0.03, -0.05])
        eps = 1e-5
        fd = (loss(scenes.compose(pose, eps * direction)) - loss(scenes.compose(pose, -eps * direction))) / (2 * eps)
        analytic = float(tangent.sum(axis=0) @ direction)
        assert abs(fd - analytic) <= 5e-5 * max(1.0, abs(fd)), (fd, analytic)
        # the 12-component packed gradient is the same derivative
        _, packed, _, _ = scene.vjp(values, seeds, pose, precision="fp64", volume=False, integration="cell_gauss")
        _, midpoint_packed, _, _ = scene.vjp(values, seeds, pose, precision="fp64", volume=False)
        assert np.isfinite(packed).all() and not np.array_equal(packed, midpoint_packed)


# ---------------------------------------------------------------------------
# Ray-coordinate derivatives (python/dpt/batched_projection.py:37-111)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("mode", [("fp64", "midpoint"), ("fp64", "cell_gauss"), ("fp32", "midpoint")])
def test_ray_coordinate_vjp_matches_central_differences(mode):
    """Each ray's 6 derivatives against central differences of its own
    optical depth when its world source or endpoint moves along one axis
    (the oblique bundle: axis-aligned, oblique and missing rays)."""
    precision, integration = mode
    grid, sources, endpoints, pose, values = scenes.oblique_scene()
    samples = 97
    scene = O.Scene(grid, samples, rays=(sources, endpoints))
    seeds = np.asarray([0.7, -0.3, 0.2, 0.6, 0.0])
    rays, status = scene.ray_vjp(values, seeds, pose, precision=precision, integration=integration)
    assert status.tolist() == [0]
    assert np.abs(rays[4]).max() == 0.0  # zero cotangent
    h = 1e-4 if precision == "fp64" else 2e-2
    tol = 2e-5 if precision == "fp64" else 5e-3
    for ray in range(4):
        for j in range(6):
            def depth(delta):
                s, e = sources.copy(), endpoints.copy()
                (s if j < 3 else e)[ray, j % 3] += delta
                image, _ = O.Scene(grid, samples, rays=(s, e)).forward(
                    values, pose, precision=precision, integration=integration)
                return float(image.reshape(-1)[ray])
            fd = seeds[ray] * (depth(h) - depth(-h)) / (2 * h)
            scale = max(1e-3, np.abs(rays[ray]).max())
            assert abs(rays[ray, j] - fd) <= tol * scale, (ray, j, rays[ray, j], fd)


@pytest.mark.parametrize("seed_scale", [1.0, 1000.0])
@pytest.mark.parametrize("samples", [2, 7, 1024])
@pytest.mark.parametrize("direction", scenes.SHARED_FACE_DIRECTIONS)
def test_fp32_volume_vjp_matches_independent_per_sample_scatter(direction, samples, seed_scale):
    """tests/python/warp_cpu/test_projection_shared_face.py:91-112 on the oracle."""
    grid, sources, endpoints, values = scenes.shared_face_scene(direction)
    seeds = (seed_scale * np.asarray([0.7, -0.3, 0.0])).astype(np.float32)
    scene = O.Scene(grid, samples, rays=(sources, endpoints))
    volume, _, _, status = scene.vjp(values, seeds, pose=False)
    expected = O.explicit_fp32_scatter(grid, list(zip(sources, endpoints)), scenes.IDENTITY12, samples, seeds)
    assert status.tolist() == [0]
    assert np.count_nonzero(expected) > 0
    np.testing.assert_allclose(volume, expected, rtol=2e-5, atol=3e-7 * seed_scale)


@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_subnormal_increments_flush_to_exact_zeros(sign):
    """tests/python/warp_cpu/test_projection_shared_face.py:138-153: a seed of
    +-1e-35 makes every increment subnormal; the GPU rule gives exact zeros
    while the unflushed products do not (the planted mistake)."""
    grid, source, target, values = scenes.subnormal_scene()
    seeds = np.asarray([sign * 1e-35], dtype=np.float32)
    volume, _, _, status = O.Scene(grid, 1024, rays=(source, target)).vjp(values, seeds, pose=False)
    assert status.tolist() == [0]
    assert np.all(volume == 0.0)
    segments = list(zip(source, target))
    np.testing.assert_array_equal(O.explicit_fp32_scatter(grid, segments, scenes.IDENTITY12, 1024, seeds), volume)
    unflushed = O.explicit_fp32_scatter(grid, segments, scenes.IDENTITY12, 1024, seeds, flush=False)
    assert np.count_nonzero(unflushed) > 0 and np.all(np.sign(unflushed[unflushed != 0]) == sign)


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_overflow_scale_seeds_stay_finite_on_an_opposing_destination(sign, reverse):
    """tests/python/warp_cpu/test_projection_shared_face.py:115-135: about
    4e38 per grid sample (beyond FLT_MAX) added to -3e38 gives a finite
    +1e38 at grid sample 1; the oracle's FP64 sums carry no overflow."""
    grid, source, target, values = scenes.overflow_scene(reverse)
    seeds = np.asarray([sign * 2e38], dtype=np.float32)
    volume, _, _, status = O.Scene(grid, 1024, rays=(source, target)).vjp(values, seeds, pose=False)
    assert status.tolist() == [0]
    expected = O.explicit_fp32_scatter(grid, list(zip(source, target)), scenes.IDENTITY12, 1024, seeds)
    np.testing.assert_allclose(volume, expected, rtol=2e-5, atol=0)
    total = volume + np.full(grid.voxels, -sign * 3e38)
    assert np.all(np.isfinite(total.astype(np.float32)))
    np.testing.assert_allclose(total[1], sign * 1e38, rtol=2e-5)
