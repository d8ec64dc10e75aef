"""CPU transport oracle pinned against known answers, closed forms and
distribution tests (no GPU).

Pins: the Random123 Philox4x32-10 known-answer vectors (one of them is the
reference's own, tests/python/test_geometry_sampling_contract.py:60-63);
an independent NumPy Philox; exact source targets in vacuum; Beer's law
exp(-mu L) for single- and two-material absorbers, cellwise and Woodcock
(test_transport_physical_interactions.py:445-520); marked-Poisson null
counts; Klein-Nishina and Rayleigh chi-square (test_transport_independent_
oracles.py:287-306, test_transport_physical_interactions.py:197-300); face
ownership rules; identity determinism.  A planted-mistake test shows the
Beer-law statistic rejects a 10 % rate error.
"""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import transport as T

from . import transport_scenes as S

GOLDEN = json.loads((Path(__file__).parent / "golden" / "closed_forms.json").read_text())


@pytest.fixture(scope="module", autouse=True)
def _built():
    T.build()


def _model(scene):
    spec, physical, labels, a, s = scene[:5]
    arrays, scalars = S.bound(spec, physical, labels, a, s)
    return T.Model(arrays, scalars)


@pytest.mark.parametrize("vector", GOLDEN["philox"], ids=lambda v: v["name"])
def test_philox_known_answers(vector):
    c = vector["counter"]
    k = vector["key"]
    identity = c[0] | (c[1] << 32)
    seed = k[0] | (k[1] << 32)
    expected = tuple(int(h, 16) for h in vector["output_hex"])
    assert T.philox_words(seed, identity, c[2], c[3]) == expected
    words = S.philox_numpy(seed, np.array([identity], dtype=np.uint64), c[2], c[3])[:, 0]
    assert tuple(int(w) for w in words) == expected


def test_numpy_philox_matches_oracle_on_random_counters():
    rng = np.random.default_rng(3)
    for _ in range(50):
        seed = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        identity = int(rng.integers(0, 2**63))
        event, domain = (int(x) for x in rng.integers(0, 2**32, 2))
        u = T.random4(seed, identity, event, domain)
        ref = S.uniforms(seed, np.array([identity], dtype=np.uint64), event, domain)[:, 0]
        np.testing.assert_array_equal(u, ref)
        assert np.all((u > 0) & (u < 1))


def test_vacuum_histories_land_on_their_source_targets():
    scene = S.homogeneous(0.0, 0.0)
    spec, source = scene[0], scene[5]
    model = _model(scene)
    n, seed = 20000, 11
    out = T.trace(model, [1.0], source, seed, 0, n)
    assert np.all(out["status"] == 0) and np.all(out["events"] == 0)
    _, _, u = S.source_rays(source, seed, np.arange(n, dtype=np.uint64))
    det = spec.detector.shape[0]
    cols = np.minimum(np.floor(source.extent_mm[0] * u[0] / 4.0).astype(int), det - 1)
    rows = np.minimum(np.floor(source.extent_mm[1] * u[1] / 4.0).astype(int), det - 1)
    expected = rows * det + cols
    assert np.mean(out["pixel"] == expected) > 0.999
    assert np.all(out["score"][out["pixel"] >= 0] == 1.0)
    assert out["histogram"].tolist() == [n, 0, 0, 0, 0, 0, 0]


def _beer_statistic(detected, p):
    return (detected - p.sum()) / math.sqrt(float((p * (1 - p)).sum()))


def test_beer_law_single_absorber_cellwise():
    mu = 0.03
    scene = S.homogeneous(mu, 0.0)
    spec, source = scene[0], scene[5]
    model = _model(scene)
    n, seed = 40000, 5
    out = T.trace(model, [1.0], source, seed, 0, n)
    origin, dirs, _ = S.source_rays(source, seed, np.arange(n, dtype=np.uint64))
    lengths = S.chord_lengths(origin, dirs, spec.grid.origin_mm, spec.grid.upper_mm)
    p = np.exp(-mu * lengths)
    detected = np.count_nonzero(out["status"] == 0)
    assert abs(_beer_statistic(detected, p)) < 5.0
    # planted mistake: a 10 % rate error is rejected by the same statistic
    assert abs(_beer_statistic(detected, np.exp(-1.1 * mu * lengths))) > 5.0
    # absorbed histories had exactly one real event; survivors none
    assert np.all(out["real_events"][out["status"] == 1] == 1)
    assert np.all(out["real_events"][out["status"] == 0] == 0)


def test_beer_law_and_null_counts_two_materials_woodcock():
    mu0, mu1, majorant = 0.01, 0.05, 0.08
    n_cells = 16
    labels = np.zeros((n_cells,) * 3, dtype=np.uint8)
    labels[:, :, n_cells // 2:] = 1  # x >= 0 is material 1
    scene = S.homogeneous([mu0, mu1], [0.0, 0.0], woodcock=majorant, labels=labels)
    spec, source = scene[0], scene[5]
    model = _model(scene)
    n, seed = 40000, 9
    out = T.trace(model, [1.0, 1.0], source, seed, 0, n)
    origin, dirs, _ = S.source_rays(source, seed, np.arange(n, dtype=np.uint64))
    total, low = S.chord_lengths(origin, dirs, spec.grid.origin_mm, spec.grid.upper_mm, split_axis=0, split_at=0.0)
    high = total - low
    p = np.exp(-(mu0 * low + mu1 * high))
    survived = out["status"] == 0
    assert abs(_beer_statistic(np.count_nonzero(survived), p)) < 5.0
    # marked Poisson: survivors' events are null marks, Poisson((M-mu) L)
    expected_nulls = (majorant - mu0) * low[survived] + (majorant - mu1) * high[survived]
    z = (out["events"][survived].sum() - expected_nulls.sum()) / math.sqrt(expected_nulls.sum())
    assert abs(z) < 7.0
    assert np.all(out["real_events"][survived] == 0)


def test_klein_nishina_cosines_and_energies():
    scene = S.homogeneous(0.0, 0.01, scattering_law="free-electron-compton")
    model = _model(scene)
    energy = 60.0
    n = 100000
    cosines = np.empty(n)
    energies = np.empty(n)
    for h in range(n):
        cosines[h], energies[h] = T.compton_cosine(model, energy, 21, h, 0)
    assert np.all(np.isfinite(cosines))
    k = energy / 510.99895069
    np.testing.assert_allclose(energies / energy, 1.0 / (1.0 + k * (1.0 - cosines)), rtol=1e-14)
    edges = np.linspace(-1, 1, 33)
    observed = np.histogram(cosines, edges)[0].astype(float)
    assert S.chi_square_p(observed, S.klein_nishina_bins(energy, edges)) > 1e-6
    # planted mistake: the Thomson (low-energy) law is rejected at 60 keV
    thomson = np.diff(edges + edges**3 / 3) / (8 / 3)
    assert S.chi_square_p(observed, thomson) < 1e-6


def test_rayleigh_cosines_follow_form_factor_law():
    from dpt.transport import materials

    spec, physical, labels, a, s, dens, source = S.chest(shape=(8, 8, 8), detector=(8, 8))
    model = _model((spec, physical, labels, a, s))
    energy, material = 30.0, 2
    n = 60000
    cosines = np.array([T.rayleigh_cosine(model, energy, material, 5, h, 0) for h in range(n)])
    assert np.all(np.isfinite(cosines))
    edges = np.linspace(-1, 1, 33)
    observed = np.histogram(cosines, edges)[0].astype(float)
    table = materials.scattering_table(materials.MATERIALS[material][1])
    assert S.chi_square_p(observed, S.rayleigh_bins(table, energy, edges)) > 1e-6


def test_cell_index_assigns_faces_downstream():
    scene = S.homogeneous(0.0, 0.0)
    model = _model(scene)
    # grid faces at -16 + 2k; the point x = 0 lies on the face between cells 7 and 8
    assert T.cell_index(model, (0.0, 1.0, 1.0), (1.0, 0.0, 0.0))[0] == 8
    assert T.cell_index(model, (0.0, 1.0, 1.0), (-1.0, 0.0, 0.0))[0] == 7
    assert T.cell_index(model, (-16.0, 1.0, 1.0), (1.0, 0.0, 0.0))[0] == 0
    assert T.cell_index(model, (16.0, 1.0, 1.0), (-1.0, 0.0, 0.0))[0] == 15
    assert T.cell_index(model, (1.0, 1.0, 1.0), (0.0, 0.0, 1.0)) == (8, 8, 8)


def test_detector_pixel_support_is_half_open():
    scene = S.homogeneous(0.0, 0.0)
    model = _model(scene)
    up = (0.0, 0.0, 1.0)
    # detector lower corner (-64, -64, 100), 32 x 32 pixels of 4 mm
    assert T.detector_pixel(model, (-64.0, -64.0, 0.0), up) == 0
    assert T.detector_pixel(model, (-60.0, -64.0, 0.0), up) == 1
    assert T.detector_pixel(model, (-64.0, -60.0, 0.0), up) == 32
    assert T.detector_pixel(model, (63.999, 63.999, 0.0), up) == 32 * 32 - 1
    assert T.detector_pixel(model, (64.0, 0.0, 0.0), up) == -1
    assert T.detector_pixel(model, (0.0, 0.0, 0.0), (0.0, 0.0, -1.0)) == -1
    assert T.detector_pixel(model, (0.0, 0.0, 150.0), up) == -1


def test_scatter_direction_keeps_polar_angle_and_unit_norm():
    rng = np.random.default_rng(1)
    for _ in range(200):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        c, phi = rng.uniform(-1, 1), rng.uniform(0, 2 * math.pi)
        out = T.scatter_direction(d, c, phi)
        assert abs(np.linalg.norm(out) - 1.0) < 1e-14
        assert abs(float(np.dot(out, d)) - c) < 1e-12


def test_history_results_depend_only_on_identity():
    spec, physical, labels, a, s, dens, source = S.chest(shape=(24, 24, 24), detector=(24, 24))
    model = _model((spec, physical, labels, a, s))
    batch = T.trace(model, dens, source, 77, 1000, 3000)
    for k in (0, 17, 1234, 2999):
        one = T.trace(model, dens, source, 77, 1000 + k, 1)
        for key in ("pixel", "score", "energy", "events", "real_events", "status"):
            assert one[key][0] == batch[key][k], key
    # chunking the identities does not change the tally
    parts = [T.trace(model, dens, source, 77, 1000 + o, c, records=False) for o, c in ((0, 1000), (1000, 2000))]
    np.testing.assert_allclose(parts[0]["tally"] + parts[1]["tally"], batch["tally"], rtol=1e-12, atol=1e-300)
    assert (parts[0]["histogram"] + parts[1]["histogram"]).tolist() == batch["histogram"].tolist()


def test_spectrum_lines_follow_their_cdf():
    spec, _, labels, a, s, source = S.homogeneous(0.0, 0.0)
    from dpt.transport import PhysicalTransport, materials

    physical = PhysicalTransport((materials.scattering_table(7),), (0.0, 0.0), spectrum=S.spectrum_two_lines())
    model = _model((spec, physical, labels, a, s))
    n, seed = 20000, 3
    out = T.trace(model, [1.0], source, seed, 0, n)
    draws = S.uniforms(seed, np.arange(n, dtype=np.uint64), 0x80000002, 0)[0]
    np.testing.assert_array_equal(out["energy"], np.where(draws < 0.25, 55.0, 65.0))


def test_chest_tally_matches_per_history_scores():
    spec, physical, labels, a, s, dens, source = S.chest(shape=(24, 24, 24), detector=(24, 24))
    model = _model((spec, physical, labels, a, s))
    n = 20000
    out = T.trace(model, dens, source, 4, 0, n)
    assert out["histogram"].sum() == n and out["bits"] == 0
    assert set(np.unique(out["status"]).tolist()) <= {0, 1}
    hit = out["pixel"] >= 0
    primary = hit & (out["real_events"] == 0)
    scatter = hit & (out["real_events"] > 0)
    assert math.isclose(out["tally"][0].sum(), out["score"][primary].sum(), rel_tol=1e-12)
    assert math.isclose(out["tally"][2].sum(), out["score"][scatter].sum(), rel_tol=1e-12)
    assert primary.sum() > 0 and scatter.sum() > 0
    # Compton only lowers energy; primaries keep the source energy
    assert np.all(out["energy"][primary] == 60.0)
    assert np.all(out["energy"][scatter] <= 60.0)
    mean, var = T.moments(out["tally"], n)
    np.testing.assert_allclose(mean[0], mean[1] + mean[2], rtol=1e-12, atol=1e-300)
    assert np.all(var >= 0)
