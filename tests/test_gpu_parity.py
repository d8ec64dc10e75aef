"""HIP projector vs the CPU oracle on identical seeded inputs (GPU).

Bars: FP32 images, volume and pose gradients within 1e-5 relative L2 of the
oracle's restatement of the reference FP32 path (per-ray quantities are
compared much tighter); FP64 within 1e-10.  Every call goes through the C ABI
(``dpt._native.Plan``) or the ``Projector`` built on it.
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
from dpt.volumes import GridSpec  # noqa: E402


def _grid_spec(grid: O.Grid) -> GridSpec:
    return GridSpec(tuple(grid.shape), tuple(grid.spacing), tuple(grid.origin), tuple(grid.orientation))


class NativeScene:
    """Drive the C ABI directly with torch-owned device buffers."""

    def __init__(self, grid: O.Grid, samples: int, panels=None, rays=None, precision="float32", storage="float32",
                 extra_needs: int = 0):
        self.grid = grid
        if panels is not None:
            packed = np.asarray([p.packed() for p in panels], dtype=np.float64)
            views, (rows, cols) = len(panels), panels[0].shape
            geo = dict(panels=packed, rays=None)
        else:
            src, end = rays
            views, rows, cols = 1, 1, len(src)
            geo = dict(panels=None, rays=(np.asarray(src, np.float64), np.asarray(end, np.float64)))
        self.shape = (views, rows, cols)
        self.precision = precision
        self.storage = storage
        self.plan = _native.Plan(
            _grid_spec(grid), views=views, rows=rows, cols=cols, samples=samples, arithmetic=precision,
            storage=storage, needs=_native.VOLUME | _native.POSE | extra_needs, device=0, **geo,
        )
        self.stream = torch.cuda.current_stream().cuda_stream

    def _field(self, values):
        f = torch.as_tensor(np.ascontiguousarray(values, dtype=np.float32)).cuda()
        if self.storage == "float32":
            return f, f
        dtype = torch.bfloat16 if self.storage == "bfloat16" else torch.float16
        half = torch.empty(f.numel() + _native.HALF_PAD_ELEMENTS, dtype=dtype, device="cuda")
        _native.convert_field(f.data_ptr(), half.data_ptr(), f.numel(), self.storage, None, self.stream)
        return half, f

    def _poses(self, pose):
        pose = scenes.IDENTITY12 if pose is None else pose
        p = np.broadcast_to(np.asarray(pose, dtype=np.float64), (self.shape[0], 12))
        return torch.as_tensor(np.ascontiguousarray(p)).cuda()

    def forward(self, values, pose=None):
        field, _ = self._field(values)
        poses = self._poses(pose)
        dtype = torch.float32 if self.precision == "float32" else torch.float64
        out = torch.empty(self.shape, dtype=dtype, device="cuda")
        self.plan.forward(field.data_ptr(), poses.data_ptr(), out.data_ptr(), self.stream)
        status = self.plan.status(self.stream)
        return out.cpu().numpy(), status

    def vjp(self, values, seeds, pose=None, *, volume=True, pose_grad=True, components=12):
        field, _ = self._field(values)
        poses = self._poses(pose)
        dtype = torch.float32 if self.precision == "float32" else torch.float64
        seed = torch.as_tensor(np.ascontiguousarray(seeds).reshape(self.shape)).to(dtype).cuda()
        gv = torch.empty(self.grid.voxels, dtype=torch.float32, device="cuda")
        gp = torch.empty((self.shape[0], components), dtype=torch.float64, device="cuda")
        flags = (_native.VOLUME if volume else 0) | (_native.POSE if pose_grad else 0)
        if components == 6:
            flags |= _native.POSE_TANGENT
        self.plan.vjp(field.data_ptr(), poses.data_ptr(), seed.data_ptr(), gv.data_ptr() if volume else None,
                      gp.data_ptr() if pose_grad else None, flags, self.stream)
        status = self.plan.status(self.stream)
        return (gv.cpu().numpy() if volume else None), (gp.cpu().numpy() if pose_grad else None), status


def _rel(a, b):
    return O.relative_l2(a, b)


def _cone(n=24, views=5, rows=20, cols=28, seed=3):
    grid = O.Grid((n, n + 2, n + 4), (1.0, 0.9, 1.1), (-11.0, -12.0, -13.0))
    panels = scenes.cone_panels(grid, views, rows, cols, sod=120.0, sdd=200.0, offset=0.3)
    return grid, panels, scenes.smooth_field(grid, seed)


SCENES = {
    "oblique": lambda: scenes.oblique_scene(),
    "workflow": lambda: scenes.workflow_scene(),
    "cone": lambda: _cone(),
}


def _scene(name, samples, precision="float32", **native):
    if name == "oblique":
        grid, src, end, pose, values = scenes.oblique_scene()
        return (grid, values, pose, O.Scene(grid, samples, rays=(src, end)),
                NativeScene(grid, samples, rays=(src, end), precision=precision, **native))
    if name == "workflow":
        grid, panels, values = scenes.workflow_scene()
        pose = scenes.compose(scenes.IDENTITY12, (0.03, -0.02, 0.01, 0.04, -0.03, 0.02))
    else:
        grid, panels, values = _cone()
        pose = scenes.compose(scenes.IDENTITY12, (0.5, -0.7, 0.3, 0.02, -0.015, 0.03))
    return (grid, values, pose, O.Scene(grid, samples, panels=panels),
            NativeScene(grid, samples, panels=panels, precision=precision, **native))


@pytest.mark.parametrize("samples", [17, 97, 256])
@pytest.mark.parametrize("name", ["oblique", "workflow", "cone"])
def test_fp32_forward_matches_oracle(name, samples):
    print(f"case fp32 forward {name} S={samples}", flush=True)
    grid, values, pose, ref, gpu = _scene(name, samples)
    expected, s_ref = ref.forward(values, pose)
    actual, s_gpu = gpu.forward(values, pose)
    assert s_gpu.tolist() == s_ref.tolist() == [0] * len(s_ref)
    assert _rel(actual, expected) <= 1e-6
    np.testing.assert_allclose(actual, expected, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("samples", [17, 97, 256])
@pytest.mark.parametrize("name", ["oblique", "workflow", "cone"])
def test_fp64_forward_matches_oracle(name, samples):
    print(f"case fp64 forward {name} S={samples}", flush=True)
    grid, values, pose, ref, gpu = _scene(name, samples, "float64")
    expected, _ = ref.forward(values, pose, precision="fp64")
    actual, status = gpu.forward(values, pose)
    assert status.tolist() == [0] * len(status)
    np.testing.assert_allclose(actual, expected, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("components", [12, 6])
@pytest.mark.parametrize("mode", ["both", "volume", "pose"])
@pytest.mark.parametrize("name", ["oblique", "workflow", "cone"])
def test_fp32_vjp_matches_oracle(name, mode, components):
    print(f"case fp32 vjp {name} {mode} C={components}", flush=True)
    grid, values, pose, ref, gpu = _scene(name, 97)
    image, _ = ref.forward(values, pose)
    rng = np.random.default_rng(7)
    seeds = rng.uniform(-1.0, 1.0, image.size).astype(np.float32)
    vol, pos = mode in ("both", "volume"), mode in ("both", "pose")
    g_ref, p_ref, _, _ = ref.vjp(values, seeds, pose, volume=vol, pose=pos, components=components)
    g_gpu, p_gpu, status = gpu.vjp(values, seeds, pose, volume=vol, pose_grad=pos, components=components)
    assert status.tolist() == [0] * len(status)
    if vol:
        assert _rel(g_gpu, g_ref) <= 1e-5
    if pos:
        assert _rel(p_gpu, p_ref) <= 1e-5
        np.testing.assert_allclose(p_gpu, p_ref, rtol=1e-4, atol=1e-6 * np.abs(p_ref).max())


@pytest.mark.parametrize("segments", [None, "1", "2", "3", "8"])
@pytest.mark.parametrize("name", ["cone", "workflow"])
@pytest.mark.parametrize("cells", [False, True])
def test_fp32_vjp_sample_segments(name, segments, cells, monkeypatch):
    """Launches of few tiles split each tile's samples into segments (by
    default here: 1024 samples, 256 per segment) for the volume adjoint; the
    pose adjoint of FP32 storage runs those launches unsegmented through the
    pipelined sample loop, so each ray's pose sums are the oracle's one
    sequential FP32 sum: the pose gradient agrees up to the FP64 order of the
    cross-ray sums, joint and pose-only, linear and cell gathers."""
    print(f"case segments {name} {segments} cells={cells}", flush=True)
    if segments is not None:
        monkeypatch.setenv("DPT_AMD_VJP_SEGMENTS", segments)
    grid, values, pose, ref, gpu = _scene(name, 1024, extra_needs=_native.GATHER_CELLS if cells else 0)
    image, _ = ref.forward(values, pose)
    seeds = np.random.default_rng(29).uniform(-1.0, 1.0, image.size).astype(np.float32)
    g_ref, p_ref, _, _ = ref.vjp(values, seeds, pose)
    for vol in (True, False):
        g_gpu, p_gpu, status = gpu.vjp(values, seeds, pose, volume=vol)
        assert status.tolist() == [0] * len(status)
        if vol:
            assert _rel(g_gpu, g_ref) <= 1e-5
        assert _rel(p_gpu, p_ref) <= 1e-12, (vol, _rel(p_gpu, p_ref))


def test_half_storage_pose_segments_stay_near_the_fp64_vjp(monkeypatch):
    """16-bit storage segments the pose too (its blocked compensated sums have
    no sequential FP32 counterpart): against the FP64 VJP of the rounded field
    it stays as close as the unsegmented launch, and as the oracle's FP32 VJP
    of that field."""
    print("case half segments", flush=True)
    grid, values, pose, ref, _ = _scene("cone", 1024, storage="bfloat16")
    rounded = torch.tensor(values, dtype=torch.float32).to(torch.bfloat16).float().numpy()
    image, _ = ref.forward(rounded, pose)
    seeds = np.random.default_rng(31).uniform(-1.0, 1.0, image.size).astype(np.float32)
    _, p64, _, _ = ref.vjp(rounded, seeds.astype(np.float64), pose, precision="fp64", volume=False)
    _, p32, _, _ = ref.vjp(rounded, seeds, pose, volume=False)
    errors = []
    for segments in ("1", "4"):
        monkeypatch.setenv("DPT_AMD_VJP_SEGMENTS", segments)
        *_, gpu = _scene("cone", 1024, storage="bfloat16")
        _, p_gpu, status = gpu.vjp(values, seeds, pose, volume=False)
        assert status.tolist() == [0] * len(status)
        errors.append(_rel(p_gpu, p64))
    assert errors[1] <= 1.5 * errors[0] + 1e-7, errors
    assert max(errors) <= 1.5 * _rel(p32, p64) + 1e-7, (errors, _rel(p32, p64))


@pytest.mark.parametrize("fallback", ["1", "0"])
@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("cap,chunk", [("512", None), ("512", "7"), (None, "1"), (None, "64"), ("8192", "16")])
def test_volume_vjp_brick_paths_match_oracle(cap, chunk, precision, fallback, monkeypatch):
    """LDS bricks of every size down to the smallest capacity (512 entries),
    where most chunks overflow: the FP32 kernel then walks the rest of the
    workgroup's samples by slab chunks (fallback "1", the default) or flushes
    the chunk to device atomics ("0"); chunk lengths from 1 to 64 samples.
    The pose sums keep their order either way."""
    print(f"case bricks cap={cap} chunk={chunk} {precision} fallback={fallback}", flush=True)
    monkeypatch.setenv("DPT_AMD_SLAB_FALLBACK", fallback)
    if cap is not None:
        monkeypatch.setenv("DPT_AMD_BRICK_CAP", cap)
    if chunk is not None:
        monkeypatch.setenv("DPT_AMD_CHUNK", chunk)
    grid, panels, values = _cone(n=24, views=5, rows=20, cols=28, seed=4)
    pose = scenes.compose(scenes.IDENTITY12, (0.5, -0.7, 0.3, 0.02, -0.015, 0.03))
    ref = O.Scene(grid, 97, panels=panels)
    gpu = NativeScene(grid, 97, panels=panels, precision=precision)
    fp = "fp32" if precision == "float32" else "fp64"
    image, _ = ref.forward(values, pose, precision=fp)
    seeds = np.random.default_rng(8).uniform(-1.0, 1.0, image.size)
    if precision == "float32":
        seeds = seeds.astype(np.float32)
    g_ref, p_ref, _, _ = ref.vjp(values, seeds, pose, precision=fp)
    g_gpu, p_gpu, status = gpu.vjp(values, seeds, pose)
    assert status.tolist() == [0] * len(status)
    assert _rel(g_gpu, g_ref) <= (1e-5 if precision == "float32" else 1e-7)
    # FP32: each ray's pose sums are the oracle's sequential FP32 sums
    assert _rel(p_gpu, p_ref) <= (1e-12 if precision == "float32" else 1e-10)
    if precision == "float32":
        g_vol, _, status = gpu.vjp(values, seeds, pose, pose_grad=False)
        assert status.tolist() == [0] * len(status)
        assert _rel(g_vol, g_ref) <= 1e-5


@pytest.mark.parametrize("storage", ["bfloat16", "float16"])
def test_half_storage_slab_fallback_matches_index_chunks(storage, monkeypatch):
    """16-bit storage (blocked compensated pose sums) through the slab
    fallback: with the smallest brick most workgroups switch to slab chunks,
    whose compensation blocks differ from the index chunks' (device-atomic
    flushes); volume and pose gradients agree to FP32 rounding."""
    print(f"case half slab fallback {storage}", flush=True)
    monkeypatch.setenv("DPT_AMD_BRICK_CAP", "512")
    grid, panels, values = _cone(n=24, views=5, rows=20, cols=28, seed=5)
    pose = scenes.compose(scenes.IDENTITY12, (0.5, -0.7, 0.3, 0.02, -0.015, 0.03))
    ref = O.Scene(grid, 97, panels=panels)
    image, _ = ref.forward(values, pose)
    seeds = np.random.default_rng(9).uniform(-1.0, 1.0, image.size).astype(np.float32)
    out = {}
    for fallback in ("1", "0"):
        monkeypatch.setenv("DPT_AMD_SLAB_FALLBACK", fallback)
        gpu = NativeScene(grid, 97, panels=panels, storage=storage)
        g, p_, status = gpu.vjp(values, seeds, pose)
        assert status.tolist() == [0] * len(status)
        out[fallback] = (g, p_)
    assert _rel(out["1"][0], out["0"][0]) <= 1e-6
    assert _rel(out["1"][1], out["0"][1]) <= 1e-6


@pytest.mark.parametrize("mode", ["both", "volume"])
@pytest.mark.parametrize("cap,chunk", [(None, None), ("512", None), (None, "1"), ("512", "7"), (None, "64")])
def test_fp64_slab_chunks_match_oracle(cap, chunk, mode, monkeypatch):
    """The FP64 volume adjoint chunked by slabs of grid samples (forced on),
    joint and volume-only, from the smallest brick (device-atomic fallback
    chunks) up, chunk lengths 1 to 64."""
    print(f"case slabs cap={cap} chunk={chunk} {mode}", flush=True)
    monkeypatch.setenv("DPT_AMD_SLAB_CHUNKS", "1")
    if cap is not None:
        monkeypatch.setenv("DPT_AMD_BRICK_CAP", cap)
    if chunk is not None:
        monkeypatch.setenv("DPT_AMD_CHUNK", chunk)
    grid, panels, values = _cone(n=24, views=5, rows=20, cols=28, seed=6)
    pose = scenes.compose(scenes.IDENTITY12, (0.5, -0.7, 0.3, 0.02, -0.015, 0.03))
    ref = O.Scene(grid, 97, panels=panels)
    gpu = NativeScene(grid, 97, panels=panels, precision="float64")
    assert gpu.plan.layout()["slab_chunks"]
    image, _ = ref.forward(values, pose, precision="fp64")
    seeds = np.random.default_rng(12).uniform(-1.0, 1.0, image.size)
    pose_grad = mode == "both"
    g_ref, p_ref, _, _ = ref.vjp(values, seeds, pose, precision="fp64", pose=pose_grad)
    g_gpu, p_gpu, status = gpu.vjp(values, seeds, pose, pose_grad=pose_grad)
    assert status.tolist() == [0] * len(status)
    assert _rel(g_gpu, g_ref) <= 1e-7
    if pose_grad:
        assert _rel(p_gpu, p_ref) <= 1e-10


@pytest.mark.parametrize("segments", [None, "1", "4"])
@pytest.mark.parametrize("cap,chunk", [(None, None), ("512", None), (None, "1"), ("512", "7"), (None, "64")])
def test_fp32_slab_chunks_match_oracle(cap, chunk, segments, monkeypatch):
    """The FP32 volume adjoint chunked by slabs of grid samples (forced on):
    volume-only and, where the launch segments, beside the pipelined pose
    adjoint; from the smallest brick (device-atomic fallback chunks) up,
    chunk lengths 1 to 64, one to four sample segments."""
    print(f"case fp32 slabs cap={cap} chunk={chunk} segments={segments}", flush=True)
    monkeypatch.setenv("DPT_AMD_SLAB_CHUNKS", "1")
    if cap is not None:
        monkeypatch.setenv("DPT_AMD_BRICK_CAP", cap)
    if chunk is not None:
        monkeypatch.setenv("DPT_AMD_CHUNK", chunk)
    if segments is not None:
        monkeypatch.setenv("DPT_AMD_VJP_SEGMENTS", segments)
    grid, panels, values = _cone(n=24, views=5, rows=20, cols=28, seed=6)
    pose = scenes.compose(scenes.IDENTITY12, (0.5, -0.7, 0.3, 0.02, -0.015, 0.03))
    ref = O.Scene(grid, 512, panels=panels)
    gpu = NativeScene(grid, 512, panels=panels)
    assert gpu.plan.layout()["slab_chunks"]
    image, _ = ref.forward(values, pose)
    seeds = np.random.default_rng(12).uniform(-1.0, 1.0, image.size).astype(np.float32)
    for pose_grad in (False, True):
        g_ref, p_ref, _, _ = ref.vjp(values, seeds, pose, pose=pose_grad)
        g_gpu, p_gpu, status = gpu.vjp(values, seeds, pose, pose_grad=pose_grad)
        assert status.tolist() == [0] * len(status)
        assert _rel(g_gpu, g_ref) <= 1e-5, (pose_grad, _rel(g_gpu, g_ref))
        if pose_grad:
            assert _rel(p_gpu, p_ref) <= 1e-12


def test_plans_choose_slab_chunks_for_a_silhouette_heavy_geometry():
    """A 512 x 512 x 139 chest grid at 0.76 x 0.76 x 2.5 mm under a 512^2
    panel at 2.67 mm (the reference's gather workload, benchmarks/gather.py)
    chunks its FP64 volume adjoint by slabs with the larger brick; the
    full-scan orbit of a 256^3 cube (360 views of 512^2, S = 1024) keeps index
    chunks and the default brick."""
    import bench

    chest = bench.CHEST_GRID
    grid = O.Grid(chest["shape"], chest["spacing_mm"], chest["origin_mm"], chest["orientation"])
    ap = bench.CHEST_AP
    panel = O.Panel(ap["source_mm"], ap["origin_mm"], ap["u"], ap["v"], ap["spacing_mm"], ap["shape"])
    layout = NativeScene(grid, 1024, panels=[panel], precision="float64").plan.layout()
    assert layout["slab_chunks"] and layout["brick_cap"] == 8192, layout
    from dpt import phantoms

    cube = phantoms.centred_grid((256, 256, 256))
    geometry, _ = phantoms.circular_geometry(cube, 360, (512, 512), sod=1000.0, sdd=1500.0)
    panels = [O.Panel(p.source_mm, p.origin_mm, p.u, p.v, p.spacing_mm, p.shape) for p in geometry.views]
    grid = O.Grid(cube.shape, cube.spacing_mm, cube.origin_mm, cube.orientation)
    layout = NativeScene(grid, 1024, panels=panels, precision="float64").plan.layout()
    assert not layout["slab_chunks"] and layout["brick_cap"] == 4096, layout


@pytest.mark.parametrize("mode", ["both", "volume", "pose"])
@pytest.mark.parametrize("name", ["oblique", "workflow", "cone"])
def test_fp64_vjp_matches_oracle(name, mode):
    print(f"case fp64 vjp {name} {mode}", flush=True)
    grid, values, pose, ref, gpu = _scene(name, 97, "float64")
    image, _ = ref.forward(values, pose, precision="fp64")
    seeds = np.random.default_rng(11).uniform(-1.0, 1.0, image.size)
    vol, pos = mode in ("both", "volume"), mode in ("both", "pose")
    g_ref, p_ref, _, _ = ref.vjp(values, seeds, pose, precision="fp64", volume=vol, pose=pos)
    g_gpu, p_gpu, status = gpu.vjp(values, seeds, pose, volume=vol, pose_grad=pos)
    assert status.tolist() == [0] * len(status)
    if vol:
        # FP32-rounded increments: GPU sums them in FP64 then rounds once to FP32
        assert _rel(g_gpu, g_ref) <= 1e-7
    if pos:
        assert _rel(p_gpu, p_ref) <= 1e-10


@pytest.mark.parametrize("views_per_block", ["1", "2", "4"])
@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_vjp_views_per_workgroup_layouts_match_oracle(views_per_block, precision, monkeypatch):
    """Adjacent views (1 degree apart) sharing one workgroup and LDS brick; an
    odd view count leaves a partial last group; rows/cols not multiples of 8."""
    print(f"case vb={views_per_block} {precision}", flush=True)
    monkeypatch.setenv("DPT_AMD_VIEWS_PER_BLOCK", views_per_block)
    grid = O.Grid((20, 22, 24), (1.0, 1.0, 1.0), (-11.5, -10.5, -9.5))
    panels = [scenes.cone_panels(grid, 360, 21, 27, sod=120.0, sdd=200.0, offset=0.1)[k] for k in range(9)]
    values = scenes.smooth_field(grid, 9)
    ref = O.Scene(grid, 97, panels=panels)
    gpu = NativeScene(grid, 97, panels=panels, precision=precision)
    fp = "fp32" if precision == "float32" else "fp64"
    image, _ = ref.forward(values, precision=fp)
    seeds = np.random.default_rng(2).uniform(-1, 1, image.size)
    if precision == "float32":
        seeds = seeds.astype(np.float32)
    g_ref, p_ref, _, _ = ref.vjp(values, seeds, precision=fp)
    g_gpu, p_gpu, status = gpu.vjp(values, seeds)
    assert status.tolist() == [0] * 9
    bar = 1e-5 if precision == "float32" else 1e-7
    assert _rel(g_gpu, g_ref) <= bar
    assert _rel(p_gpu, p_ref) <= (1e-5 if precision == "float32" else 1e-10)


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_closed_form_unit_voxel_through_the_c_abi(precision):
    """tests/python/warp_cpu/test_projection_fp32.py:248-257."""
    grid = O.Grid((1, 1, 1), (1.0, 1.0, 1.0))
    rays = (np.array([[-2.0, 0.0, 0.0]]), np.array([[0.2, 0.0, 0.0]]))
    gpu = NativeScene(grid, 1024, rays=rays, precision=precision)
    image, status = gpu.forward([0.025])
    assert status.tolist() == [0]
    np.testing.assert_allclose(image.ravel(), [0.0175], rtol=2e-5)
    volume, tangent, status = gpu.vjp([0.025], [1.0], components=6)
    np.testing.assert_allclose(volume, [0.7], rtol=2e-5)
    np.testing.assert_allclose(tangent.ravel(), [-0.025, 0, 0, 0, 0, 0], rtol=2e-5, atol=1e-8)


def test_config1_forward_and_volume_vjp_match_oracle():
    """64^3 phantom, AP + lateral, 128^2 at 0.8 mm, 256 samples, central-64 mean loss."""
    from dpt import phantoms

    grid_spec = phantoms.centred_grid((64, 64, 64))
    values = phantoms.chest_phantom(grid_spec, seed=1)
    grid = O.Grid(grid_spec.shape, grid_spec.spacing_mm, grid_spec.origin_mm)
    angles = phantoms.standard_angles()
    panels = []
    for name in ("ap", "lateral"):
        a = angles[name]
        c, s = math.cos(a), math.sin(a)
        rows = cols = 128
        pitch = 0.8
        origin = (-0.5 * (cols - 1) * pitch, 500.0, -0.5 * (rows - 1) * pitch)
        rot = lambda p: (c * p[0] - s * p[1], s * p[0] + c * p[1], p[2])  # noqa: E731
        panels.append(O.Panel(rot((0.0, -1000.0, 0.0)), rot(origin), (c, s, 0.0), (0.0, 0.0, 1.0), (pitch, pitch), (rows, cols)))
    ref = O.Scene(grid, 256, panels=panels)
    gpu = NativeScene(grid, 256, panels=panels)
    expected, _ = ref.forward(values)
    actual, status = gpu.forward(values)
    assert status.tolist() == [0, 0]
    assert _rel(actual, expected) <= 1e-6
    seeds = np.zeros((2, 128, 128), dtype=np.float32)
    seeds[:, 32:96, 32:96] = 1.0 / (2 * 64 * 64)
    g_ref, p_ref, _, _ = ref.vjp(values, seeds)
    g_gpu, p_gpu, status = gpu.vjp(values, seeds)
    assert status.tolist() == [0, 0]
    assert _rel(g_gpu, g_ref) <= 1e-5
    assert _rel(p_gpu, p_ref) <= 1e-5


@pytest.mark.parametrize("storage", ["bfloat16", "float16"])
def test_half_storage_forward_and_gradients_within_budget_of_fp64(storage):
    grid, panels, values = _cone(n=32, views=6, rows=24, cols=32)
    ref = O.Scene(grid, 256, panels=panels)
    gpu = NativeScene(grid, 256, panels=panels, storage=storage)
    expected, _ = ref.forward(values, precision="fp64")
    actual, status = gpu.forward(values)
    assert status.tolist() == [0] * 6
    seeds = np.random.default_rng(5).uniform(-1, 1, expected.size)
    g_ref, p_ref, _, _ = ref.vjp(values, seeds, precision="fp64")
    g_gpu, p_gpu, status = gpu.vjp(values, seeds.astype(np.float32))
    errors = (_rel(actual, expected), _rel(g_gpu, g_ref), _rel(p_gpu, p_ref))
    print(f"half storage {storage}: image {errors[0]:.3e} volume {errors[1]:.3e} pose {errors[2]:.3e}", flush=True)
    # measured on MI355X: bf16 1.9e-4 / 4.2e-6 / 2.0e-3, f16 2.5e-5 / 4.2e-6 /
    # 2.8e-4 (image / volume / pose, relative L2 against the FP64 oracle);
    # the bars leave a factor of 2.5
    bars = {"bfloat16": (5e-4, 1e-5, 5e-3), "float16": (6e-5, 1e-5, 7e-4)}[storage]
    assert all(e <= b for e, b in zip(errors, bars)), (errors, bars)
