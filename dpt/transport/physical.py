"""Atomic scattering tables, source spectra and the physical-transport binding.

Host-side mirror of python/dpt/transport/physical.py:37-262 and
transport/source_spectrum.py:23-100 for the simulated path: Rayleigh
scattering by the independent-atom law (1+c^2) F(q)^2 with
q = E sqrt((1-c)/2) / hc, a fixed Woodcock majorant M(E) and a discrete
source spectrum.  Fluorescence lines and bound-electron Compton are not part
of this path (only the empty fluorescence record is accepted).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

from .model import TransportError, _physical

HC_KEV_ANGSTROM = 12.398419843320026


@dataclass(frozen=True)
class AtomicScatteringTable:
    """Form factor F(q) and incoherent function S(q) on q nodes (physical.py:37-103).

    q in inverse angstroms increases strictly from zero; F(0) = Z, S(0) = 0,
    0 <= F, S <= Z; both interpolate linearly in q with no extension.
    """

    atomic_number: int
    momentum_transfer_inverse_angstrom: tuple[float, ...]
    form_factor: tuple[float, ...]
    scattering_function: tuple[float, ...]
    provenance: str

    def __post_init__(self) -> None:
        if type(self.atomic_number) is not int or not 1 <= self.atomic_number <= 100:
            raise TransportError("atomic number must be in 1..100")
        for name in ("momentum_transfer_inverse_angstrom", "form_factor", "scattering_function"):
            object.__setattr__(self, name, tuple(float(x) for x in getattr(self, name)))
        q = self.momentum_transfer_inverse_angstrom
        if len(q) < 2 or q[0] != 0 or any(b <= a for a, b in zip(q, q[1:])):
            raise TransportError("atomic scattering q nodes must increase strictly from zero")
        if len(self.form_factor) != len(q) or len(self.scattering_function) != len(q):
            raise TransportError("atomic scattering values must match q nodes")
        for value in q:
            _physical(value, "momentum transfer", nonnegative=True)
        for name, values in (("form factor", self.form_factor), ("scattering function", self.scattering_function)):
            for value in values:
                _physical(value, name, nonnegative=True)
                if value > self.atomic_number:
                    raise TransportError(f"{name} exceeds the atomic number")
        if self.form_factor[0] != self.atomic_number or self.scattering_function[0] != 0:
            raise TransportError("atomic scattering requires F(0)=Z and S(0)=0")
        if type(self.provenance) is not str or not self.provenance.strip():
            raise TransportError("atomic scattering requires data provenance")


@dataclass(frozen=True)
class SourceSpectrum:
    """Discrete emitted-energy law P(E = E_k) = p_k (source_spectrum.py:23-100)."""

    energy_kev: tuple[float, ...]
    probabilities: tuple[float, ...]
    provenance: str

    def __post_init__(self) -> None:
        object.__setattr__(self, "energy_kev", tuple(float(e) for e in self.energy_kev))
        object.__setattr__(self, "probabilities", tuple(float(p) for p in self.probabilities))
        if not self.energy_kev or len(self.energy_kev) != len(self.probabilities):
            raise TransportError("source spectrum needs matching nonempty energies and masses")
        for energy in self.energy_kev:
            _physical(energy, "spectrum energy", positive=True)
        if any(b <= a for a, b in zip(self.energy_kev, self.energy_kev[1:])):
            raise TransportError("source spectrum energies must increase strictly")
        for p in self.probabilities:
            _physical(p, "spectrum probability", positive=True)
        if abs(math.fsum(self.probabilities) - 1.0) > 2e-15:
            raise TransportError("source spectrum probabilities must sum to one")
        cdf = self.cdf
        if any(b <= a for a, b in zip((0.0, *cdf[:-1]), cdf)):
            raise TransportError("source spectrum masses must remain distinguishable in binary64 CDF")
        if type(self.provenance) is not str or not self.provenance.strip():
            raise TransportError("source spectrum requires provenance")

    @property
    def cdf(self) -> tuple[float, ...]:
        """C_k = fsum(p_0..p_k), the final value exactly one (source_spectrum.py:77-91)."""
        n = len(self.probabilities)
        return tuple(math.fsum(self.probabilities[: i + 1]) if i + 1 < n else 1.0 for i in range(n))


@dataclass(frozen=True)
class PhysicalTransport:
    """Rayleigh partials, majorant, spectrum and cutoff (physical.py:165-262).

    rayleigh_mm_inverse: (M*K,) coherent partial of scattering, material-major,
    energy fastest, mm^-1 per unit density.  majorant_mm_inverse: () or (K,);
    when present every occurring material's rate must stay strictly below it.
    """

    scattering_tables: tuple[AtomicScatteringTable, ...]
    rayleigh_mm_inverse: tuple[float, ...]
    fluorescence: tuple = ()
    cutoff_kev: float = 0.0
    majorant_mm_inverse: tuple[float, ...] = ()
    spectrum: SourceSpectrum | None = None

    def __post_init__(self) -> None:
        object.__setattr__(self, "scattering_tables", tuple(self.scattering_tables))
        object.__setattr__(self, "rayleigh_mm_inverse", tuple(float(x) for x in self.rayleigh_mm_inverse))
        object.__setattr__(self, "majorant_mm_inverse", tuple(float(x) for x in self.majorant_mm_inverse))
        object.__setattr__(self, "fluorescence", tuple(tuple(lines) for lines in self.fluorescence))
        _physical(self.cutoff_kev, "energy cutoff", nonnegative=True)
        if not self.scattering_tables or any(not isinstance(t, AtomicScatteringTable) for t in self.scattering_tables):
            raise TransportError("physical transport requires one atomic scattering table per material")
        for value in self.rayleigh_mm_inverse:
            _physical(value, "Rayleigh coefficient", nonnegative=True)
        for value in self.majorant_mm_inverse:
            _physical(value, "Woodcock majorant", positive=True)
        if any(len(lines) for lines in self.fluorescence):
            raise TransportError("fluorescence lines are not supported by this transport path")
        if self.spectrum is not None and not isinstance(self.spectrum, SourceSpectrum):
            raise TransportError("physical spectrum must be a SourceSpectrum record")
