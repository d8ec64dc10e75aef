"""The measured workloads at their own sizes (GPU).

* Full scan (256^3 FP32 chest phantom, 360-view circular orbit, 512^2 panel,
  1024 samples): four of the 360 views (0, 90, 180, 270) through the C ABI
  against the CPU oracle on the same field and residual cotangent -- image
  bitwise, volume and pose gradients within 1e-5 relative L2 -- and one view
  in FP64 sampling (1e-12 image, 1e-7 volume with its FP32-rounded
  increments, 1e-10 pose).
* Precision form (BF16 / FP16 storage, FP32 sampling, compensated sums) of
  the same scan at S = 512 and 1024 on every tenth view, against the GPU
  FP64 path at the same field, pose and cotangent.
* Serving: pose fits on a 512 x 507 x 76 volume at 0.45 x 0.45 x 2.5 mm and a
  128^3 volume with 512^2 / S = 1024, 256^2 / S = 512 and 128^2 / S = 256
  requests: bit-identical alone and mixed, every pose recovered.
* Monte Carlo: 128^3 labels, 256^2 tallies, 1e7 histories against the oracle
  history by history, and across seeds per pixel.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt import _native, phantoms, serving, transport  # noqa: E402
from dpt.geometry import ConeBeamIntrinsics, RigidTransform, compose_pose  # noqa: E402
from dpt.torch import Projector  # noqa: E402
from dpt.transport.simulation import bind_model  # noqa: E402
from oracle import transport as T  # noqa: E402

from .test_gpu_parity import NativeScene  # noqa: E402

SHAPE, VIEWS, DETECTOR, SAMPLES = (256, 256, 256), 360, (512, 512), 1024


def _full_scan():
    grid = phantoms.centred_grid(SHAPE)
    values = phantoms.chest_phantom(grid, seed=2026)
    geometry, _ = phantoms.circular_geometry(grid, VIEWS, DETECTOR, sod=1000.0, sdd=1500.0)
    return grid, values, geometry


def _oracle_grid(grid):
    return O.Grid(grid.shape, grid.spacing_mm, grid.origin_mm, grid.orientation)


def _panels(geometry, picks):
    return [O.Panel(v.source_mm, v.origin_mm, v.u, v.v, v.spacing_mm, v.shape)
            for v in (geometry.views[k] for k in picks)]


def _rel(a, b):
    return O.relative_l2(a, b)


def test_full_scan_views_match_the_oracle():
    print("case full scan 4 views FP32", flush=True)
    grid, values, geometry = _full_scan()
    panels = _panels(geometry, (0, 90, 180, 270))
    ref = O.Scene(_oracle_grid(grid), SAMPLES, panels=panels)
    gpu = NativeScene(_oracle_grid(grid), SAMPLES, panels=panels, extra_needs=_native.GATHER_CELLS)
    expected, status = ref.forward(values)
    actual, gstatus = gpu.forward(values)
    assert status.tolist() == gstatus.tolist() == [0] * 4
    assert np.array_equal(actual.view(np.int32), expected.view(np.int32))
    seeds = (expected - 0.5 * expected.mean()).astype(np.float32)
    g_ref, p_ref, _, _ = ref.vjp(values, seeds)
    g_gpu, p_gpu, gstatus = gpu.vjp(values, seeds)
    assert gstatus.tolist() == [0] * 4
    volume, pose = _rel(g_gpu, g_ref), _rel(p_gpu, p_ref)
    print(f"full scan FP32: volume {volume:.2e}, pose {pose:.2e}", flush=True)
    assert volume <= 1e-5 and pose <= 1e-5


def test_full_scan_view_in_fp64_matches_the_oracle():
    print("case full scan 1 view FP64", flush=True)
    grid, values, geometry = _full_scan()
    panels = _panels(geometry, (45,))
    ref = O.Scene(_oracle_grid(grid), SAMPLES, panels=panels)
    gpu = NativeScene(_oracle_grid(grid), SAMPLES, panels=panels, precision="float64")
    expected, _ = ref.forward(values, precision="fp64")
    actual, status = gpu.forward(values)
    assert status.tolist() == [0]
    np.testing.assert_allclose(actual, expected, rtol=1e-12, atol=1e-15)
    seeds = expected - 0.5 * expected.mean()
    g_ref, p_ref, _, _ = ref.vjp(values, seeds, precision="fp64")
    g_gpu, p_gpu, status = gpu.vjp(values, seeds)
    assert status.tolist() == [0]
    assert _rel(g_gpu, g_ref) <= 1e-7
    assert _rel(p_gpu, p_ref) <= 1e-10


@pytest.mark.parametrize("samples", [512, 1024])
def test_precision_form_at_full_size_against_the_fp64_path(samples):
    print(f"case precision form S={samples}", flush=True)
    grid, values, full = _full_scan()
    geometry = type(full)(full.views[::10])
    device = torch.device("cuda:0")
    field = torch.tensor(values, device=device, requires_grad=True)
    pose = torch.zeros(6, dtype=torch.float64, device=device, requires_grad=True)
    rng = np.random.default_rng(7)
    xi = np.concatenate([rng.uniform(-5, 5, 3), np.deg2rad(rng.uniform(-2, 2, 3))])
    with torch.no_grad():
        target = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)(
            field.detach(), torch.tensor(xi, dtype=torch.float64, device=device))
    image64 = Projector(grid=grid, geometry=geometry, precision="fp64", samples_per_ray=samples)(field, pose)
    cotangent = (image64 - target.double()).detach()
    g64, p64 = torch.autograd.grad(image64, (field, pose), cotangent)

    def rel(a, b):
        a, b = a.detach().double().reshape(-1), b.detach().double().reshape(-1)
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    # measured on MI355X (image / volume / pose relative L2 against FP64):
    # S = 512: BF16 1.1e-4 / 1.0e-5 / 7.0e-5, FP16 1.4e-5 / 1.0e-5 / 7.7e-5,
    # FP32 7.4e-7 / 1.0e-5 / 7.2e-5; S = 1024: 1.1e-4 / 8.3e-6 / 8.0e-5,
    # 1.4e-5 / 8.3e-6 / 2.5e-5, 8.0e-7 / 8.3e-6 / 1.9e-5 (the pose error is the
    # FP32 sampling's, whatever the storage); the bars leave a factor 5-10
    bars = {"bfloat16": (1e-3, 5e-5, 5e-4), "float16": (1.5e-4, 5e-5, 5e-4), "float32": (1e-5, 5e-5, 5e-4)}
    for storage, bar in bars.items():
        module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples, storage=storage)
        image = module(field, pose)
        g, p = torch.autograd.grad(image, (field, pose), cotangent.float())
        errors = (rel(image, image64), rel(g, g64), rel(p, p64))
        print(f"precision S={samples} {storage}: image {errors[0]:.2e} volume {errors[1]:.2e} "
              f"pose {errors[2]:.2e}", flush=True)
        assert all(e <= b for e, b in zip(errors, bar)), (storage, errors, bar)


def _fit_request(key, volume, grid, values, det, samples, views, iterations, seed):
    rng = np.random.default_rng(seed)
    angles = rng.uniform(0.0, 2 * math.pi, views)
    xi = np.concatenate([rng.uniform(-5, 5, 3), np.radians(rng.uniform(-3, 3, 3))])
    nx, ny, nz = grid.xyz
    diag = math.hypot(nx * grid.spacing_mm[0], ny * grid.spacing_mm[1], nz * grid.spacing_mm[2])
    sod = max(1000.0, 2.0 * diag)
    pitch = phantoms.coverage_pitch(grid, sod, 1.5 * sod, angles, (det, det), margin=1.15)
    geometry = ConeBeamIntrinsics(sod, 1.5 * sod, (det, det), (pitch, pitch), centre_mm=grid.centre_mm()).batch(angles)
    truth = compose_pose(RigidTransform(), tuple(xi)).packed()
    with torch.no_grad():
        target = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)(
            torch.tensor(values, device="cuda:0"), packed_pose=torch.tensor(truth, dtype=torch.float64, device="cuda:0"))
    return serving.FitRequest(key, volume, geometry, target.cpu().numpy(), samples, iterations), truth


def test_serving_slice_is_batch_invariant_and_converges():
    print("case serving slice", flush=True)
    brain = phantoms.centred_grid((76, 507, 512), (0.45, 0.45, 2.5))
    small = phantoms.centred_grid((128, 128, 128))
    volumes = {"brain": (brain, phantoms.chest_phantom(brain, seed=31)),
               "small": (small, phantoms.chest_phantom(small, seed=32))}
    specs = [("a", "brain", 512, 1024, 4, 30, 1), ("b", "small", 256, 512, 3, 30, 2),
             ("c", "brain", 128, 256, 2, 40, 3), ("d", "small", 512, 768, 2, 30, 4)]
    requests, truths = [], {}
    for key, vol, det, samples, views, iterations, seed in specs:
        grid, values = volumes[vol]
        request, truth = _fit_request(key, vol, grid, values, det, samples, views, iterations, seed)
        requests.append(request)
        truths[key] = truth
    registry = serving.VolumeRegistry()
    for name, (grid, values) in volumes.items():
        registry.add(name, values, grid, storage="bfloat16", layout="cells")
    runs = {}
    for name, batch in (("mixed", requests), ("alone", requests[:1])):
        engine = serving.PoseFitEngine(registry, max_requests=8, max_tiles=1 << 17)
        runs[name] = {r.key: r for r in engine.run(batch, realtime=False).results}
        engine.close()
    mixed, alone = runs["mixed"], runs["alone"]
    assert np.array_equal(alone["a"].pose, mixed["a"].pose) and alone["a"].loss == mixed["a"].loss
    for key, result in mixed.items():
        angle, shift = serving.pose_error(result.pose, truths[key])
        print(f"serving {key}: rotation {angle:.2e} deg, translation {shift:.2e} mm", flush=True)
        assert angle < 0.05 and shift < 0.05, (key, angle, shift)


def test_monte_carlo_at_full_size_matches_the_oracle():
    print("case monte carlo 1e7 histories", flush=True)
    spec, physical, labels, a, s, dens, source = transport.chest_transport_scene()
    assert labels.shape == (128, 128, 128) and spec.detector.image_shape == (256, 256)
    n = 10_000_000
    ws = transport.prepare_transport(spec, material_ids=labels, absorption=a, scattering=s, max_histories=n,
                                     physical=physical)
    model = T.Model(*bind_model(spec, material_ids=labels, absorption=a, scattering=s, physical=physical))
    gpu = transport.trace_histories(dens, source=source, batch=transport.HistoryBatch(2026, 0, n), workspace=ws)
    ref = T.trace(model, dens, source, 2026, 0, n)
    same = np.ones(n, dtype=bool)
    for key in ("pixel", "events", "real_events", "status"):
        same &= gpu[key].cpu().numpy() == ref[key]
    print(f"monte carlo: identical histories {same.mean():.6f}", flush=True)
    assert same.mean() >= 0.999
    score = gpu["score"].cpu().numpy()
    np.testing.assert_array_equal(score[same], ref["score"][same])
    # tallies of the identical histories agree to rounding
    pixel = ref["pixel"][same]
    hit = pixel >= 0
    pixels = int(np.prod(spec.detector.image_shape))
    t_gpu = np.bincount(pixel[hit], weights=score[same][hit], minlength=pixels)
    t_ref = np.bincount(pixel[hit], weights=ref["score"][same][hit], minlength=pixels)
    np.testing.assert_allclose(t_gpu, t_ref, rtol=1e-9, atol=0)
    del gpu, ref
    # independent seeds: per-pixel z and total fluence
    sim = transport.prepare_simulation(ws)
    result = transport.simulate(dens, source=source, batch=transport.HistoryBatch(77, 0, n), workspace=sim)
    cpu = T.trace(model, dens, source, 78, 0, n, records=False)
    mean, var = T.moments(cpu["tally"], n)
    m_gpu, v_gpu = result.mean.cpu().numpy(), result.estimated_variance_of_mean.cpu().numpy()
    ok = (m_gpu + mean[0]) * n >= 100
    z = (m_gpu[ok] - mean[0][ok]) / np.sqrt(v_gpu[ok] + var[0][ok])
    print(f"monte carlo: z mean {z.mean():.3f} std {z.std():.3f} over {z.size} pixels", flush=True)
    assert abs(z.mean()) < 4.0 / math.sqrt(z.size) and 0.8 < z.std() < 1.2
    total_sigma = math.sqrt(float(v_gpu.sum() + var[0].sum()))
    assert abs(float(m_gpu.sum() - mean[0].sum())) < 3.0 * total_sigma
