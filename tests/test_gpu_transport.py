"""Monte Carlo transport on the GPU against the CPU oracle.

Per-history parity: the HIP walk and the oracle draw the same counter-based
Philox streams, so the same history identity must take the same path; the
test requires identical pixel / events / real events / status and bitwise
equal score and energy for at least 99.9 % of histories (transcendental
library calls may differ in the last bit, which can flip a rare decision).
Statistical parity: tallies from independent seeds agree per pixel in
z = (m_gpu - m_cpu) / sqrt(v_gpu + v_cpu) and in total fluence within 3 sigma.
Error paths are forced from host data (budgets, energy support, majorant).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dpt import transport  # noqa: E402
from oracle import transport as T  # noqa: E402

from . import transport_scenes as S  # noqa: E402

FIELDS = ("pixel", "events", "real_events", "status")


def _scene(case):
    if case == "chest_mono":
        return S.chest()
    if case == "chest_spectrum":
        return S.chest(spectrum=True)
    if case == "chest_packed":
        return S.chest(packed=True)
    if case == "absorber_cellwise":
        scene = S.homogeneous(0.03, 0.0)
        return (*scene[:5], np.array([1.0]), scene[5])
    if case == "two_materials_woodcock":
        labels = np.zeros((16,) * 3, dtype=np.uint8)
        labels[:, :, 8:] = 1
        scene = S.homogeneous([0.01, 0.05], [0.02, 0.01], woodcock=0.1, labels=labels,
                              scattering_law="free-electron-compton", energies=(10.0, 40.0, 70.0))
        return (*scene[:5], np.array([1.0, 1.0]), scene[5])
    if case == "isotropic_cellwise":
        scene = S.homogeneous(0.01, 0.04)
        return (*scene[:5], np.array([1.0]), scene[5])
    raise KeyError(case)


def _blocks(image, shape, b):
    rows, cols = shape
    return image.reshape(rows // b, b, cols // b, b).sum(axis=(1, 3)).ravel()


def _workspace(scene, capacity):
    spec, physical, labels, a, s = scene[:5]
    return transport.prepare_transport(spec, material_ids=labels, absorption=a, scattering=s,
                                       max_histories=capacity, physical=physical)


def _oracle_model(scene):
    spec, physical, labels, a, s = scene[:5]
    return T.Model(*S.bound(spec, physical, labels, a, s))


@pytest.mark.parametrize(
    "case", ["chest_mono", "chest_spectrum", "chest_packed", "absorber_cellwise", "two_materials_woodcock",
             "isotropic_cellwise"])
def test_histories_match_oracle(case):
    print(f"[transport parity] {case}", flush=True)
    scene = _scene(case)
    dens, source = scene[5], scene[6]
    n, seed, first = 60000, 12345, 1000
    ws = _workspace(scene, n)
    gpu = transport.trace_histories(dens, source=source, batch=transport.HistoryBatch(seed, first, n), workspace=ws)
    ref = T.trace(_oracle_model(scene), dens, source, seed, first, n)
    g = {k: gpu[k].cpu().numpy() for k in (*FIELDS, "score", "energy")}
    same = np.ones(n, dtype=bool)
    for key in FIELDS:
        same &= g[key] == ref[key]
    assert same.mean() >= 0.999, (case, same.mean())
    np.testing.assert_array_equal(g["score"][same], ref["score"][same])
    np.testing.assert_array_equal(g["energy"][same], ref["energy"][same])
    assert sum(gpu["histogram"]) == n
    assert np.abs(np.asarray(gpu["histogram"]) - ref["histogram"]).sum() <= 2 * (n - same.sum())


def test_simulate_matches_oracle_tallies_for_the_same_histories():
    scene = S.chest()
    spec, dens, source = scene[0], scene[5], scene[6]
    n, seed = 200000, 99
    ws = _workspace(scene, n)
    sim = transport.prepare_simulation(ws)
    result = transport.simulate(dens, source=source, batch=transport.HistoryBatch(seed, 0, n), workspace=sim)
    ref = T.trace(_oracle_model(scene), dens, source, seed, 0, n, records=False)
    mean, var = T.moments(ref["tally"], n)
    assert result.status_histogram.total == n
    assert list(result.status_histogram.counts) == ref["histogram"].tolist()
    for k, (m_img, v_img) in enumerate(((result.mean, result.estimated_variance_of_mean),
                                         (result.primary_mean, result.primary_variance_of_mean),
                                         (result.scatter_mean, result.scatter_variance_of_mean))):
        np.testing.assert_allclose(m_img.cpu().numpy(), mean[k], rtol=1e-9, atol=1e-18)
        np.testing.assert_allclose(v_img.cpu().numpy(), var[k], rtol=1e-6, atol=1e-18)
    assert result.image_shape == spec.detector.image_shape


@pytest.mark.parametrize("spectrum", [False, True])
def test_tallies_agree_statistically_across_seeds(spectrum):
    print(f"[transport statistics] spectrum={spectrum}", flush=True)
    scene = S.chest(shape=(64, 64, 64), detector=(32, 32), spectrum=spectrum)
    dens, source = scene[5], scene[6]
    n_gpu, n_cpu = 8_000_000, 2_000_000
    ws = _workspace(scene, n_gpu)
    sim = transport.prepare_simulation(ws)
    result = transport.simulate(dens, source=source, batch=transport.HistoryBatch(2024, 0, n_gpu), workspace=sim)
    ref = T.trace(_oracle_model(scene), dens, source, 7, 0, n_cpu, records=False)
    mean, var = T.moments(ref["tally"], n_cpu)
    images = ((result.mean, result.estimated_variance_of_mean),
              (result.primary_mean, result.primary_variance_of_mean),
              (result.scatter_mean, result.scatter_variance_of_mean))
    side = scene[0].detector.image_shape
    for k, (m_img, v_img) in enumerate(images):
        m_gpu, v_gpu = m_img.cpu().numpy(), v_img.cpu().numpy()
        # photon-count scores of weight 1: hits = mean * N.  The normal
        # approximation of z needs enough hits in both estimates, so sparse
        # images (scatter) are compared on b x b pixel blocks; block sums of
        # per-pixel variances bound the block variance (exclusive hits).
        # Blocks are selected on the pooled estimate, which does not bias z.
        for b in (1, 2, 4, 8):
            mg, vg, mc, vc = (_blocks(x, side, b) for x in (m_gpu, v_gpu, mean[k], var[k]))
            pooled = (mg * n_gpu + mc * n_cpu) / (n_gpu + n_cpu)
            ok = pooled * n_cpu >= 50
            if ok.mean() > 0.5:
                break
        assert ok.mean() > 0.5, (k, b, ok.mean())
        z = (mg[ok] - mc[ok]) / np.sqrt(vg[ok] + vc[ok])
        assert abs(z.mean()) < 4.0 / math.sqrt(z.size) and 0.7 < z.std() < 1.3, (k, b, z.mean(), z.std())
        assert np.max(np.abs(z)) < 6.0
        # total fluence within 3 sigma; per-history scores land in one pixel,
        # so summing per-pixel variances bounds the variance of the total
        total_sigma = math.sqrt(float(v_gpu.sum() + var[k].sum()))
        assert abs(float(m_gpu.sum() - mean[k].sum())) < 3.0 * total_sigma


def test_chunked_trace_equals_single_launch():
    scene = S.chest()
    dens, source = scene[5], scene[6]
    n = 100000
    ws = _workspace(scene, n)
    sim = transport.prepare_simulation(ws)
    batch = transport.HistoryBatch(5, 0, n)
    one = transport.simulate(dens, source=source, batch=batch, workspace=sim)
    m1 = one.mean.cpu().numpy().copy()
    h1 = one.status_histogram.counts
    two = transport.simulate(dens, source=source, batch=batch, workspace=sim, chunk=30000)
    np.testing.assert_allclose(two.mean.cpu().numpy(), m1, rtol=1e-12, atol=0)
    assert two.status_histogram.counts == h1


def test_guards_raise_typed_errors():
    # energy support: Compton scattering below the lowest table node
    scene = S.homogeneous(0.0, 0.05, energies=(59.0, 61.0), scattering_law="free-electron-compton")
    ws = _workspace(scene, 1000)
    sim = transport.prepare_simulation(ws)
    with pytest.raises(transport.TransportError, match="energy support"):
        transport.simulate([1.0], source=scene[5], batch=transport.HistoryBatch(1, 0, 1000), workspace=sim)
    # event budget
    scene = S.homogeneous(0.0, 0.5, max_events=2)
    ws = _workspace(scene, 1000)
    sim = transport.prepare_simulation(ws)
    with pytest.raises(transport.IncompleteHistoryError):
        transport.simulate([1.0], source=scene[5], batch=transport.HistoryBatch(1, 0, 1000), workspace=sim)
    # majorant reached by a density scale
    scene = _scene("two_materials_woodcock")
    ws = _workspace(scene, 1000)
    sim = transport.prepare_simulation(ws)
    with pytest.raises(transport.TransportError, match="majorant"):
        transport.simulate([1.0, 2.0], source=scene[6], batch=transport.HistoryBatch(1, 0, 1000), workspace=sim)
    # capacity
    with pytest.raises(transport.TransportError, match="2..workspace"):
        transport.simulate([1.0, 1.0], source=scene[6], batch=transport.HistoryBatch(1, 0, 1001), workspace=sim)
    # the workspace keeps working after the errors
    result = transport.simulate([1.0, 1.0], source=scene[6], batch=transport.HistoryBatch(1, 0, 1000), workspace=sim)
    assert result.status_histogram.total == 1000


def test_object_pose_moves_the_phantom_not_the_detector():
    from dpt.geometry import RigidTransform

    scene = S.homogeneous(0.03, 0.0, pose=RigidTransform(translation_mm=(0.0, 0.0, 20.0)))
    ws = _workspace(scene, 20000)
    out = transport.trace_histories([1.0], source=scene[5], batch=transport.HistoryBatch(3, 0, 20000), workspace=ws)
    ref = T.trace(_oracle_model(scene), [1.0], scene[5], 3, 0, 20000)
    for key in FIELDS:
        assert np.mean(out[key].cpu().numpy() == ref[key]) > 0.999
