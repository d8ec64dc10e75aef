"""Small transport scenes shared by the CPU oracle tests and the GPU parity tests."""

from __future__ import annotations

import math

import numpy as np

from dpt.geometry import DetectorGeometry, RigidTransform
from dpt.transport import (
    AtomicScatteringTable,
    MaterialGrid,
    OrientedDetector,
    PhysicalTransport,
    PointConeSource,
    SourceSpectrum,
    TransportSpec,
)
from dpt.transport import materials
from dpt.transport.simulation import bind_model

NAMESPACE = 0x80000000


def box_geometry(n=16, spacing=2.0, det=32, pitch=4.0, source_z=-200.0, detector_z=100.0):
    """Cube of n^3 cells centred on the origin; source below, detector above."""
    half = n * spacing / 2
    grid = MaterialGrid((-half, -half, -half), (spacing,) * 3, (n, n, n), 1)
    corner = -det * pitch / 2 + pitch / 2
    geometry = DetectorGeometry((0.0, 0.0, source_z), (corner, corner, detector_z), (1.0, 0.0, 0.0),
                                (0.0, 1.0, 0.0), (pitch, pitch), (det, det))
    return grid, geometry


def homogeneous(mu_a, mu_s, *, energies=(50.0, 70.0), energy=60.0, woodcock=None, n=16, spacing=2.0,
                scattering_law="isotropic-elastic", materials_count=1, labels=None, det=32, pitch=4.0,
                max_events=4096, pose=None):
    """One- or two-material box with energy-independent rates (mm^-1).

    mu_a, mu_s: per-material sequences (or scalars).  woodcock: majorant
    (mm^-1) or None for the cellwise analogue walk without physics tables."""
    mu_a = np.atleast_1d(np.asarray(mu_a, dtype=np.float64))
    mu_s = np.atleast_1d(np.asarray(mu_s, dtype=np.float64))
    grid, geometry = box_geometry(n, spacing, det, pitch)
    grid = MaterialGrid(grid.origin_mm, grid.spacing_mm, grid.shape, len(mu_a))
    K = len(energies)
    absorption = np.repeat(mu_a, K)
    scattering = np.repeat(mu_s, K)
    physical = None
    if woodcock is not None:
        tables = tuple(materials.scattering_table(7) for _ in mu_a)
        physical = PhysicalTransport(tables, tuple(np.zeros(len(mu_a) * K)), majorant_mm_inverse=(woodcock,) * K)
    spec = TransportSpec(grid, OrientedDetector(geometry), energy, "test: constant rates", tuple(energies),
                         scattering_law=scattering_law, max_events=max_events,
                         object_pose=RigidTransform() if pose is None else pose)
    if labels is None:
        labels = np.zeros(grid.shape, dtype=np.uint8)
    return spec, physical, labels, absorption, scattering, PointConeSource(geometry, 1.0)


def bound(spec, physical, labels, absorption, scattering):
    return bind_model(spec, material_ids=labels, absorption=absorption, scattering=scattering, physical=physical)


def chest(shape=(48, 48, 48), detector=(48, 48), spectrum=False, seed=7, packed=False):
    spec, physical, labels, a, s, dens, source = materials.chest_transport_scene(
        seed=seed, shape=shape, detector=detector, spectrum=spectrum)
    if packed:
        spec = spec.with_execution(compton_sampler="packed-uniform-rejection")
    return spec, physical, labels, a, s, dens, source


# ---------------------------------------------------------------- numpy Philox
def philox_numpy(seed, identity, event, domain):
    """Vectorised Philox4x32-10 over identity arrays (independent of the C oracle)."""
    identity = np.asarray(identity, dtype=np.uint64)
    mask = np.uint64(0xFFFFFFFF)
    c0 = identity & mask
    c1 = identity >> np.uint64(32)
    c2 = np.full_like(c0, np.uint64(event))
    c3 = np.full_like(c0, np.uint64(domain))
    k0 = np.uint64(seed & 0xFFFFFFFF)
    k1 = np.uint64(seed >> 32)
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c0
        p1 = np.uint64(0xCD9E8D57) * c2
        c0, c1, c2, c3 = ((p1 >> np.uint64(32)) ^ c1 ^ k0) & mask, p1 & mask, ((p0 >> np.uint64(32)) ^ c3 ^ k1) & mask, p0 & mask
        k0 = (k0 + np.uint64(0x9E3779B9)) & mask
        k1 = (k1 + np.uint64(0xBB67AE85)) & mask
    return np.stack([c0, c1, c2, c3]).astype(np.uint32)


def uniforms(seed, identity, event, domain):
    return (philox_numpy(seed, identity, event, domain).astype(np.float64) + 0.5) / 4294967296.0


def source_rays(source: PointConeSource, seed: int, histories):
    """Positions and unit directions of the point-cone law (numpy restatement)."""
    u = uniforms(seed, histories, NAMESPACE, 0)
    g = source.geometry
    corner = np.asarray(source.corner_mm)
    end = corner[None, :] + (source.extent_mm[0] * u[0])[:, None] * np.asarray(g.u)[None, :] \
        + (source.extent_mm[1] * u[1])[:, None] * np.asarray(g.v)[None, :]
    delta = end - np.asarray(g.source_mm)[None, :]
    direction = delta / np.linalg.norm(delta, axis=1, keepdims=True)
    return np.asarray(g.source_mm), direction, u


def chord_lengths(origin, directions, lower, upper, split_axis=None, split_at=None):
    """Slab-method chord of each ray through the box; optionally the part on
    the low side of the plane x_axis = split_at."""
    t0 = np.zeros(len(directions))
    t1 = np.full(len(directions), np.inf)
    for a in range(3):
        d = directions[:, a]
        with np.errstate(divide="ignore"):
            ta = (lower[a] - origin[a]) / d
            tb = (upper[a] - origin[a]) / d
        t0 = np.maximum(t0, np.minimum(ta, tb))
        t1 = np.minimum(t1, np.maximum(ta, tb))
    length = np.maximum(t1 - t0, 0.0)
    if split_axis is None:
        return length
    d = directions[:, split_axis]
    with np.errstate(divide="ignore"):
        ts = (split_at - origin[split_axis]) / d
    # portion of [t0, t1] with coordinate below split_at (d > 0: before ts)
    low_end = np.clip(np.where(d > 0, ts, t1), t0, t1)
    low_start = np.clip(np.where(d > 0, t0, ts), t0, t1)
    low = np.where(length > 0, np.maximum(low_end - low_start, 0.0), 0.0)
    return length, low


def klein_nishina_bins(energy_kev, edges):
    """Probability of each cosine bin under the Klein-Nishina density."""
    k = energy_kev / 510.99895069
    fine = np.linspace(-1.0, 1.0, 200001)
    r = 1.0 / (1.0 + k * (1.0 - fine))
    density = r * r * (r + 1.0 / r - (1.0 - fine * fine))
    cdf = np.concatenate([[0.0], np.cumsum((density[1:] + density[:-1]) * 0.5 * np.diff(fine))])
    cdf /= cdf[-1]
    return np.diff(np.interp(edges, fine, cdf))


def rayleigh_bins(table: AtomicScatteringTable, energy_kev, edges):
    fine = np.linspace(-1.0, 1.0, 200001)
    q = energy_kev / 12.398419843320026 * np.sqrt((1.0 - fine) * 0.5)
    f = np.interp(q, table.momentum_transfer_inverse_angstrom, table.form_factor)
    density = (1.0 + fine * fine) * f * f
    cdf = np.concatenate([[0.0], np.cumsum((density[1:] + density[:-1]) * 0.5 * np.diff(fine))])
    cdf /= cdf[-1]
    return np.diff(np.interp(edges, fine, cdf))


def chi_square_p(observed, probabilities):
    from scipy import stats

    expected = probabilities * observed.sum()
    chi2 = float(((observed - expected) ** 2 / expected).sum())
    return float(stats.chi2.sf(chi2, len(observed) - 1))


def spectrum_two_lines():
    return SourceSpectrum((55.0, 65.0), (0.25, 0.75), "test: two lines")


def unit_sphere_angle(a, b):
    return math.degrees(math.acos(max(-1.0, min(1.0, float(np.dot(a, b))))))
