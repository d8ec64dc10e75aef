"""Host-side contracts (no GPU): ABI exports, status decoding, geometry,
Projector argument rules, fake-tensor tracing and phantoms."""

from __future__ import annotations

import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest
import torch

import dpt
from dpt import _native, phantoms
from dpt.contracts import ContractError, NumericalError, TrialDomainError, raise_view_statuses, status_error
from dpt.geometry import BatchedGeometry, ConeBeamIntrinsics, DetectorGeometry, RigidTransform, compose_pose
from dpt.projection import ProjectionSpec, configuration, decode
from dpt.torch import Projector
from dpt.torch import compose_pose as torch_compose
from dpt.volumes import GridSpec
from tests import scenes

HEADER = Path(__file__).resolve().parent.parent / "include" / "dpt_amd.h"


def _declared_functions() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(dpt_[a-z_0-9]+)\s*\(", text, re.M)))


def test_header_declares_the_entry_points_the_binding_uses():
    assert _declared_functions() == sorted(_native.EXPORTS)


def test_library_loads_and_exports_every_declared_symbol():
    lib = _native.lib()
    for name in _declared_functions():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value
    assert lib.dpt_abi_version() == 1


def test_library_rejects_invalid_plans_without_a_device():
    lib = _native.lib()
    out = ctypes.c_void_p()
    rc = lib.dpt_plan_create(None, 1, 1, 1, None, None, None, 1, 0, 0, 0, 0, ctypes.byref(out))
    assert rc == -1 and b"grid" in lib.dpt_last_error()
    desc = _native.GridDesc()
    desc.nx = desc.ny = desc.nz = 4
    rc = lib.dpt_plan_create(ctypes.byref(desc), 1, 1, 1, None, None, None, 0, 0, 0, 0, 0, ctypes.byref(out))
    assert rc == -1 and b"samples" in lib.dpt_last_error()
    rc = lib.dpt_plan_create(ctypes.byref(desc), 1, 1, 1, None, None, None, 8, 7, 0, 0, 0, ctypes.byref(out))
    assert rc == -1 and b"arithmetic" in lib.dpt_last_error()
    # gather indices use 24-bit multiplies: ny * nz must not exceed 2^24
    desc.nx, desc.ny, desc.nz = 2, 4097, 4097
    rc = lib.dpt_plan_create(ctypes.byref(desc), 1, 1, 1, None, None, None, 8, 0, 0, 0, 0, ctypes.byref(out))
    assert rc == -1 and b"2^24" in lib.dpt_last_error()


def test_cell_layout_counts_and_argument_checks_without_a_device():
    assert _native.cell_count((5, 6, 7)) == 4 * 5 * 6
    assert _native.cell_count((1, 1, 9)) == 8
    assert _native.cell_count((1, 1, 1)) == 1
    lib = _native.lib()
    assert lib.dpt_cell_count(0, 1, 1) == -1
    assert lib.dpt_convert_cells(None, None, 4, 4, 4, 0, None) == -1
    assert b"src and dst" in lib.dpt_last_error()


@pytest.mark.parametrize(
    "code, kind, message",
    [
        (1, ContractError, "invalid rigid pose"),
        (8, TrialDomainError, "finite and non-negative"),
        (2, NumericalError, "non-finite"),
        (4, NumericalError, "numerical execution failure"),
        (1 | 2 | 8, ContractError, "invalid rigid pose"),
        (2 | 8, TrialDomainError, "finite and non-negative"),
        (32, NumericalError, "unknown device status"),
    ],
)
def test_status_bits_map_to_reference_exceptions(code, kind, message):
    error = status_error(code, "projection")
    assert type(error) is kind and message in str(error)


def test_first_failing_view_is_raised_with_its_index():
    with pytest.raises(ContractError, match=r"projection view 2.*invalid rigid"):
        raise_view_statuses([0, 0, 1, 2], "projection")
    raise_view_statuses([0, 0, 0], "projection")


def test_cone_intrinsics_follow_reference_conventions():
    intr = ConeBeamIntrinsics(1000.0, 1500.0, (4, 6), (0.5, 0.25))
    view = intr.view(0.0)
    assert view.source_mm == pytest.approx((0.0, -1000.0, 0.0))
    # panel centre at (0, SDD-SOD, 0); pixel (0,0) offset by half the extent
    assert view.origin_mm == pytest.approx((-0.5 * 5 * 0.5, 500.0, -0.5 * 3 * 0.25))
    assert view.u == pytest.approx((1.0, 0.0, 0.0)) and view.v == pytest.approx((0.0, 0.0, 1.0))
    quarter = intr.view(math.pi / 2)
    assert quarter.source_mm == pytest.approx((1000.0, 0.0, 0.0), abs=1e-9)
    assert quarter.u == pytest.approx((0.0, 1.0, 0.0), abs=1e-12)
    assert view.pixel_centre(3, 5) == pytest.approx((-1.25 + 5 * 0.5, 500.0, -0.375 + 3 * 0.25))


def test_host_and_torch_compose_pose_agree_and_act_on_the_right():
    anchor = RigidTransform((0.0, -1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0), (1.0, 2.0, 3.0))
    xi = (0.3, -0.2, 0.1, 0.05, -0.07, 0.02)
    host = compose_pose(anchor, xi).packed()
    torch_value = torch_compose(torch.tensor(anchor.packed(), dtype=torch.float64), torch.tensor(xi, dtype=torch.float64))
    np.testing.assert_allclose(torch_value.numpy(), host, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(host, scenes.compose(anchor.packed(), xi), rtol=1e-13, atol=1e-13)
    # quarter-turn anchor maps a local x displacement onto world y (geometry.py:752)
    moved = compose_pose(anchor, (1.0, 0, 0, 0, 0, 0)).translation_mm
    assert moved == pytest.approx((1.0, 3.0, 3.0))


def test_torch_pose_chart_passes_gradcheck():
    """The exp-map chart's autograd (matrix_exp) against finite differences,
    as the reference pins it (tests/python/test_projector_workflow.py:96-98)."""
    anchor = torch.tensor(RigidTransform((0.0, -1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0), (1.0, 2.0, 3.0)).packed(),
                          dtype=torch.float64)
    xi = torch.tensor([0.3, -0.2, 0.1, 0.05, -0.07, 0.02], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: torch_compose(anchor, x), (xi,), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_grid_object_to_grid_matrix_is_diag_inverse_spacing_times_orientation_transpose():
    c, s = math.cos(0.3), math.sin(0.3)
    grid = GridSpec((2, 3, 4), (0.5, 2.0, 4.0), (1.0, 2.0, 3.0), (c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0))
    g = np.asarray(grid.object_to_grid_matrix()).reshape(3, 3)
    o = np.asarray(grid.orientation).reshape(3, 3)
    np.testing.assert_allclose(g, np.diag([2.0, 0.5, 0.25]) @ o.T, rtol=1e-15)
    point = (2.0, -1.0, 5.0)
    np.testing.assert_allclose(g @ (np.asarray(point) - grid.origin_mm), grid.object_to_grid(point), rtol=1e-14)


def test_grid_and_detector_contracts():
    with pytest.raises(ContractError):
        GridSpec((0, 1, 1), (1.0, 1.0, 1.0))
    with pytest.raises(ContractError):
        GridSpec((1, 1, 1), (1.0, -1.0, 1.0))
    with pytest.raises(ContractError, match="orthonormal"):
        GridSpec((1, 1, 1), (1.0, 1.0, 1.0), orientation=(2.0, 0, 0, 0, 1, 0, 0, 0, 1))
    with pytest.raises(ContractError, match="unit length"):
        DetectorGeometry((0, -5, 0), (0, 5, 0), (2.0, 0, 0), (0, 0, 1.0), (1.0, 1.0), (2, 2))
    with pytest.raises(ContractError, match="outside the detector plane"):
        DetectorGeometry((0, 5, 0), (0, 5, 0), (1.0, 0, 0), (0, 0, 1.0), (1.0, 1.0), (2, 2))


def _small_projector(**kwargs):
    grid = phantoms.centred_grid((8, 9, 10))
    geometry, _ = phantoms.circular_geometry(grid, 3, (6, 7), sod=100.0, sdd=150.0)
    return Projector(grid=grid, geometry=geometry, **kwargs), grid, geometry


def test_projector_argument_rules():
    with pytest.raises(ContractError, match="fp32 or fp64"):
        _small_projector(precision="fp16")
    with pytest.raises(ContractError, match="choose samples_per_ray or step_mm"):
        _small_projector(samples_per_ray=8, step_mm=1.0)
    module, grid, _ = _small_projector(step_mm=0.5)
    diagonal = math.hypot(10, 9, 8)
    assert module.samples_per_ray == math.ceil(diagonal / 0.5)
    module, _, _ = _small_projector()
    assert module.samples_per_ray == 1024
    with pytest.raises(ContractError, match="half storage requires FP32"):
        _small_projector(precision="fp64", storage="bfloat16")


def test_projector_rejects_bad_fields_and_poses_on_the_host():
    module, grid, _ = _small_projector(samples_per_ray=16)
    with pytest.raises(ContractError, match="grid shape"):
        module(torch.zeros((8, 9, 11)))
    with pytest.raises(ContractError, match="float32"):
        module(torch.zeros(grid.shape, dtype=torch.float64))
    field = torch.zeros(grid.shape)
    with pytest.raises(ContractError, match="float64"):
        module(field, torch.zeros(6))
    with pytest.raises(ContractError, match="shape"):
        module(field, torch.zeros(5, dtype=torch.float64))
    with pytest.raises(ContractError, match="choose pose or packed_pose"):
        module(field, torch.zeros(6, dtype=torch.float64), packed_pose=torch.zeros(12, dtype=torch.float64))


def test_cpu_tensors_are_refused_there_is_no_cpu_execution_path():
    module, grid, _ = _small_projector(samples_per_ray=16)
    with pytest.raises(ContractError, match="ROCm"):
        module(torch.zeros(grid.shape))


def test_fake_tensor_tracing_gives_output_and_gradient_metadata():
    from torch._subclasses.fake_tensor import FakeTensorMode

    module, grid, geometry = _small_projector(samples_per_ray=16)
    with FakeTensorMode():
        field = torch.empty(grid.shape, requires_grad=True)
        pose = torch.zeros(6, dtype=torch.float64, requires_grad=True)
        image = module(field, pose)
        assert tuple(image.shape) == geometry.shape and image.dtype == torch.float32
    meta = module(torch.empty(grid.shape, device="meta"))
    assert meta.shape == geometry.shape
    module64, _, _ = _small_projector(samples_per_ray=16, precision="fp64")
    assert module64(torch.empty(grid.shape, device="meta")).dtype == torch.float64


def test_custom_op_schemas_match_the_reference_operators():
    from dpt.torch import _projection_ops  # noqa: F401

    project = torch.ops.dpt.project.default._schema
    assert [a.name for a in project.arguments] == ["field", "poses", "value", "active_volume", "active_pose"]
    vjp = torch.ops.dpt.project_vjp.default._schema
    assert [a.name for a in vjp.arguments] == ["field", "poses", "cotangent", "value", "active_volume", "active_pose"]
    out = torch.ops.dpt.project_out.default._schema
    assert [a.name for a in out.arguments if a.alias_info and a.alias_info.is_write] == ["out"]


def test_configuration_round_trips():
    grid = phantoms.centred_grid((4, 5, 6), (0.5, 0.7, 0.9))
    geometry, _ = phantoms.circular_geometry(grid, 4, (3, 5), sod=50.0, sdd=80.0)
    spec = ProjectionSpec(33, "float32", "bfloat16")
    decoded = decode(configuration(grid, geometry, spec))
    assert decoded.grid == grid and decoded.geometry == geometry and decoded.spec == spec
    assert decoded.image_shape == (4, 3, 5)


def test_phantom_is_seeded_non_negative_and_labelled():
    grid = phantoms.centred_grid((24, 24, 24))
    a = phantoms.chest_phantom(grid, seed=3)
    b = phantoms.chest_phantom(grid, seed=3)
    c = phantoms.chest_phantom(grid, seed=4)
    assert a.dtype == np.float32 and np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0.0
    labels = phantoms.chest_labels(a)
    assert set(np.unique(labels)) <= {0, 1, 2} and (labels == 2).any() and (labels == 0).any()
    # body ~0.020, lungs ~0.005, bone ~0.050 mm^-1 present
    assert np.isclose(np.median(a[a > 0.015]), 0.020, atol=1e-3)


def test_sampled_ellipsoid_and_analytic_chord():
    grid = GridSpec((3, 3, 3), (1.0, 1.0, 1.0), (-1.0, -1.0, -1.0))
    values = phantoms.sample_ellipsoids(grid, (phantoms.Ellipsoid((1.0, 1.0, 1.0), 0.25),))
    assert values.shape == (3, 3, 3) and int((values > 0).sum()) == 7  # phantoms.py:114-120
    assert phantoms.Ellipsoid((2.0, 1.0, 1.0), 0.25).optical_depth((-3.0, 0, 0), (0.0, 0, 0)) == pytest.approx(0.5)


def test_coverage_pitch_covers_every_support_corner():
    grid = phantoms.centred_grid((16, 16, 16))
    geometry, pitch = phantoms.circular_geometry(grid, 8, (32, 32), sod=100.0, sdd=150.0)
    for panel in geometry.views:
        src = np.asarray(panel.source_mm)
        u, v = np.asarray(panel.u), np.asarray(panel.v)
        centre = np.asarray(panel.origin_mm) + (31 * pitch * u + 31 * pitch * v) / 2
        normal = (centre - src) / np.linalg.norm(centre - src)
        for corner in grid.support_corners_mm():
            x = np.asarray(corner)
            hit = src + (x - src) * (np.linalg.norm(centre - src) / ((x - src) @ normal))
            assert abs((hit - centre) @ u) <= 16 * pitch and abs((hit - centre) @ v) <= 16 * pitch


def test_api_project_argument_rules():
    grid = phantoms.centred_grid((4, 4, 4))
    with pytest.raises(ContractError, match="prepared"):
        dpt.project(torch.zeros(grid.shape))
    with pytest.raises(ContractError, match="geometry is required"):
        dpt.project(dpt.ImagingVolume(np.zeros(grid.shape), grid), device="meta")
    with pytest.raises(ContractError, match="ImagingVolume"):
        dpt.project(np.zeros(grid.shape))


def test_batched_geometry_requires_equal_shapes():
    a = ConeBeamIntrinsics(100.0, 150.0, (4, 4), (1.0, 1.0)).view(0.0)
    b = ConeBeamIntrinsics(100.0, 150.0, (4, 5), (1.0, 1.0)).view(0.0)
    with pytest.raises(ContractError, match="share"):
        BatchedGeometry((a, b))


def test_parallel_and_cylindrical_panels_validate_and_round_trip():
    """geometry.py:306-431 contracts, and the JSON configuration keeps each
    view's kind."""
    import math

    from dpt.geometry import BatchedGeometry, CurvedDetectorGeometry, DetectorGeometry, ParallelGeometry
    from dpt.projection import ProjectionSpec, configuration, decode
    from dpt.volumes import GridSpec

    parallel = ParallelGeometry((5.0, 0.0, 0.0), (1.0, 0.0, 0.0), 10.0, (0.0, 1.0, 0.0), (0.0, 0.0, 1.0),
                                (0.5, 0.5), (3, 4))
    source, end = parallel.segment(1, 2)
    assert end == (5.0, 1.0, 0.5) and source == (-5.0, 1.0, 0.5)
    with pytest.raises(ContractError, match="unit length"):
        ParallelGeometry((5.0, 0.0, 0.0), (2.0, 0.0, 0.0), 10.0, (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.5, 0.5), (3, 4))
    curved = CurvedDetectorGeometry((0.0, 0.0, 0.0), (100.0, 0.0, -1.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0),
                                    (2.0, 1.0), (3, 5))
    assert math.isclose(curved.radius_mm, 100.0)
    for col in range(5):
        p = curved.pixel_centre(2, col)
        assert math.isclose(math.hypot(p[0], p[1]), 100.0, rel_tol=1e-14)
        assert math.isclose(math.atan2(p[1], p[0]), col * 2.0 / 100.0, rel_tol=1e-12, abs_tol=1e-15)
        assert p[2] == 1.0
    with pytest.raises(ContractError, match="tangent"):
        CurvedDetectorGeometry((0.0, 0.0, 0.0), (100.0, 0.0, 0.0), (0.6, 0.8, 0.0), (0.0, 0.0, 1.0), (2.0, 1.0), (3, 5))
    with pytest.raises(ContractError, match="revolution"):
        CurvedDetectorGeometry((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (2.0, 1.0), (3, 5))
    cone = DetectorGeometry((-50.0, 0.0, 0.0), (30.0, -1.0, -0.5), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.5, 0.5),
                            (3, 4))
    batch = BatchedGeometry((cone, parallel, CurvedDetectorGeometry((0.0, 0.0, 0.0), (100.0, 0.0, -1.0),
                                                                    (0.0, 1.0, 0.0), (0.0, 0.0, 1.0),
                                                                    (2.0, 1.0), (3, 4))))
    assert not batch.flat_cone and batch.shape == (3, 3, 4)
    value = configuration(GridSpec((4, 5, 6), (1.0, 1.0, 1.0)), batch, ProjectionSpec(16, "float64"))
    decoded = decode(value)
    assert decoded.geometry == batch
    assert [type(v).__name__ for v in decoded.geometry.views] == [
        "DetectorGeometry", "ParallelGeometry", "CurvedDetectorGeometry"]
