"""Built-in three-material interaction tables and the chest transport scene.

The reference ships no material cross-section tables
(python/dpt/transport/__init__.py:3-4: "no material coefficients ... are
bundled"); its users supply tables from a provider such as xraylib.  This
module builds a self-consistent synthetic set, stated here in full:

* incoherent scattering: free-electron Klein-Nishina total cross-section per
  electron (closed form, Klein & Nishina 1929; series 1 - 2k + 26k^2/5 -
  133k^3/10 below k = 0.02) times electrons per gram N_A Z/A, matching the
  free-electron Compton law the walk samples;
* coherent scattering: the Thomson-weighted integral of the same form factor
  the walk samples, sigma_R = pi r_e^2 int_{-1}^{1} (1+c^2) F(q)^2 dc per
  effective atom, with a two-shell hydrogen-like F(q) (form_factor_model),
  so rate and angular law agree;
* photoelectric absorption: tau/rho = tau60 * (60 keV / E)^3 anchored at
  60 keV to round magnitudes of published tables (air 0.0155, soft tissue
  0.0185, cortical bone 0.125 cm^2/g); absorption edges are ignored.

Tables hold mass coefficients / 10 (cm^2/g -> mm^-1 per g/cm^3): the rate
is density * table with densities 0.0012 / 1.0 / 1.85 g/cm^3.  These are
benchmark inputs, not authoritative physical data.
"""

from __future__ import annotations

import math

import numpy as np

from .. import phantoms
from ..geometry import ConeBeamIntrinsics
from .model import MaterialGrid, OrientedDetector, TransportSpec
from .physical import HC_KEV_ANGSTROM, AtomicScatteringTable, PhysicalTransport, SourceSpectrum
from .source import PointConeSource

AVOGADRO = 6.02214076e23
ELECTRON_RADIUS_CM = 2.8179403262e-13
ELECTRON_REST_KEV = 510.99895069

PROVENANCE = (
    "synthetic: free-electron Klein-Nishina incoherent, two-shell hydrogen-like F(q) coherent, "
    "(60/E)^3 photoelectric anchored at 60 keV (air 0.0155, soft tissue 0.0185, bone 0.125 cm^2/g)"
)

# name, effective Z, Z/A, photoelectric tau/rho at 60 keV (cm^2/g), nominal density g/cm^3
MATERIALS = (
    ("air", 7, 0.49919, 0.0155, 0.0012),
    ("soft tissue", 7, 0.55000, 0.0185, 1.0),
    ("cortical bone", 13, 0.51478, 0.125, 1.85),
)
DENSITIES = tuple(m[4] for m in MATERIALS)

ENERGY_NODES_KEV = tuple([float(e) for e in range(1, 11)] + [float(e) for e in range(12, 151, 2)])
Q_NODES = tuple([0.0] + list(np.geomspace(0.01, 13.0, 63)))


def klein_nishina_cross_section(energy_kev: float) -> float:
    """Total Klein-Nishina cross-section per free electron, cm^2."""
    k = energy_kev / ELECTRON_REST_KEV
    thomson = 8.0 * math.pi * ELECTRON_RADIUS_CM**2 / 3.0
    if k < 0.02:
        return thomson * (1.0 - 2.0 * k + 5.2 * k * k - 13.3 * k**3)
    a = 1.0 + 2.0 * k
    log = math.log(a)
    bracket = (1.0 + k) / k**2 * (2.0 * (1.0 + k) / a - log / k) + log / (2.0 * k) - (1.0 + 3.0 * k) / a**2
    return 2.0 * math.pi * ELECTRON_RADIUS_CM**2 * bracket


def form_factor_model(z: int, q) -> np.ndarray:
    """Two hydrogen-like shells: 2 K electrons (screened charge Z - 0.3,
    q_K = (Z - 0.3) / (2 pi a0)) and Z - 2 outer electrons (q_V = 0.15 Z^(1/3)),
    each contributing n / (1 + (q/q_s)^2)^2 (q = sin(theta/2)/lambda)."""
    q = np.asarray(q, dtype=np.float64)
    q_k = (z - 0.3) / (2.0 * math.pi * 0.529177)
    q_v = 0.15 * z ** (1.0 / 3.0)
    inner = min(z, 2)
    return inner / (1.0 + (q / q_k) ** 2) ** 2 + (z - inner) / (1.0 + (q / q_v) ** 2) ** 2


def scattering_table(z: int) -> AtomicScatteringTable:
    f = form_factor_model(z, Q_NODES)
    f[0] = float(z)
    s = np.clip(z - f * f / z, 0.0, float(z))
    s[0] = 0.0
    return AtomicScatteringTable(z, Q_NODES, tuple(f), tuple(s), PROVENANCE)


def coherent_cross_section(z: int, energy_kev: float, nodes: int = 4001) -> float:
    """pi r_e^2 int_{-1}^{1} (1+c^2) F(q)^2 dc per effective atom (cm^2), with
    F linearly interpolated on Q_NODES as the walk evaluates it (Simpson)."""
    c = np.linspace(-1.0, 1.0, nodes)
    q = energy_kev / HC_KEV_ANGSTROM * np.sqrt((1.0 - c) * 0.5)
    table = scattering_table(z)
    f = np.interp(q, table.momentum_transfer_inverse_angstrom, table.form_factor)
    g = (1.0 + c * c) * f * f
    h = c[1] - c[0]
    integral = h / 3.0 * (g[0] + g[-1] + 4.0 * g[1:-1:2].sum() + 2.0 * g[2:-1:2].sum())
    return math.pi * ELECTRON_RADIUS_CM**2 * integral


def coefficient_tables(energies=ENERGY_NODES_KEV):
    """(absorption, scattering, rayleigh) as (M*K,) mm^-1 per g/cm^3 tables."""
    a, s, r = [], [], []
    for _, z, z_over_a, tau60, _ in MATERIALS:
        electrons = AVOGADRO * z_over_a
        for e in energies:
            incoherent = klein_nishina_cross_section(e) * electrons
            coherent = coherent_cross_section(z, e) * electrons / z
            a.append(tau60 * (60.0 / e) ** 3 / 10.0)
            s.append((incoherent + coherent) / 10.0)
            r.append(coherent / 10.0)
    return np.asarray(a), np.asarray(s), np.asarray(r)


def majorant_table(absorption, scattering, densities=DENSITIES, slack: float = 1e-3, energies=ENERGY_NODES_KEV):
    """M(E_k) = (1 + slack) * max_m density_m * (a + s)_m(E_k)."""
    k = len(energies)
    rates = [densities[m] * (absorption[m * k:(m + 1) * k] + scattering[m * k:(m + 1) * k])
             for m in range(len(densities))]
    return tuple(float(x) for x in (1.0 + slack) * np.max(np.stack(rates), axis=0))


def spectrum_40_120() -> SourceSpectrum:
    """81 lines at 40..120 keV (1 keV bins): a synthetic filtered Kramers shape
    p(E) ~ (121 - E)/E * exp(-3 (30/E)^3), normalised with fsum."""
    energies = [float(e) for e in range(40, 121)]
    w = [(121.0 - e) / e * math.exp(-3.0 * (30.0 / e) ** 3) for e in energies]
    total = math.fsum(w)
    return SourceSpectrum(tuple(energies), tuple(x / total for x in w), "synthetic filtered Kramers 40-120 keV")


def chest_transport_scene(*, seed: int = 2026, shape=(128, 128, 128), detector=(256, 256), spectrum: bool = False,
                          energy_kev: float = 60.0, view_angle: float = math.pi, sod: float = 1000.0,
                          sdd: float = 1500.0, max_events: int = 4096, max_angle_trials: int = 65536):
    """Label map of the seeded chest phantom (0 air, 1 soft tissue, 2 bone) at
    1 mm, AP cone-beam point source, primary/scatter tallies on ``detector``.

    The Rayleigh rejection sampler accepts about 0.16 % of proposals at
    120 keV, so 4096 trials would exhaust with probability ~1e-3 per event;
    the scene raises the angular budget to the 65536 trial domains the
    physical path reserves (probability ~e^-105).

    Returns (spec, physical, labels (nz,ny,nx) uint8, absorption, scattering,
    densities, source)."""
    grid = phantoms.centred_grid(shape)
    labels = phantoms.chest_labels(phantoms.chest_phantom(grid, seed=seed))
    pitch = phantoms.coverage_pitch(grid, sod, sdd, [view_angle], detector, margin=1.05)
    geometry = ConeBeamIntrinsics(sod, sdd, detector, (pitch, pitch), centre_mm=grid.centre_mm()).view(view_angle)
    a, s, r = coefficient_tables()
    lines = spectrum_40_120() if spectrum else None
    physical = PhysicalTransport(
        tuple(scattering_table(m[1]) for m in MATERIALS),
        tuple(r),
        majorant_mm_inverse=majorant_table(a, s),
        spectrum=lines,
    )
    spec = TransportSpec(
        MaterialGrid.from_grid_spec(grid, materials=len(MATERIALS)),
        OrientedDetector(geometry),
        120.0 if spectrum else energy_kev,
        PROVENANCE,
        ENERGY_NODES_KEV,
        scattering_law="free-electron-compton",
        max_events=max_events,
        max_angle_trials=max_angle_trials,
    )
    return spec, physical, labels, a, s, np.asarray(DENSITIES), PointConeSource(geometry, 1.0)
