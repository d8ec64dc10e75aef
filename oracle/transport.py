"""Python binding of the CPU transport oracle (TEST INFRASTRUCTURE).

Only tests, the smoke check and bench.py's CPU-baseline leg may import this
module, and only as the checker.  It loads ``_build/libtransport_ref.so``
(built from ``transport_ref.c`` by the Makefile) and runs the reference walk
(python/dpt/transport/kernels.py:941-1563) in binary64 on the host.

Models are passed as the ``(arrays, scalars)`` pair that the product's
``bind_model`` produces, so the oracle and the GPU run the same bound model.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "libtransport_ref.so"
_LIB = None


class _Model(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("materials", ctypes.c_int32),
        ("energy_bins", ctypes.c_int32),
        ("spectrum_bins", ctypes.c_int32),
        ("max_events", ctypes.c_int32),
        ("max_crossings", ctypes.c_int32),
        ("max_angle_trials", ctypes.c_int32),
        ("compton", ctypes.c_int32),
        ("compton_packed", ctypes.c_int32),
        ("energy_score", ctypes.c_int32),
        ("physics", ctypes.c_int32),
        ("woodcock", ctypes.c_int32),
        ("det_shape", ctypes.c_int32 * 2),
        ("origin", ctypes.c_double * 3),
        ("spacing", ctypes.c_double * 3),
        ("det_corner", ctypes.c_double * 3),
        ("det_u", ctypes.c_double * 3),
        ("det_v", ctypes.c_double * 3),
        ("det_normal", ctypes.c_double * 3),
        ("det_spacing", ctypes.c_double * 2),
        ("inv_rot", ctypes.c_double * 9),
        ("inv_trans", ctypes.c_double * 3),
        ("source_energy", ctypes.c_double),
        ("cutoff", ctypes.c_double),
        ("material_ids", ctypes.c_void_p),
        ("energies", ctypes.c_void_p),
        ("absorption", ctypes.c_void_p),
        ("scattering", ctypes.c_void_p),
        ("rayleigh", ctypes.c_void_p),
        ("majorant", ctypes.c_void_p),
        ("majorant_max_density", ctypes.c_void_p),
        ("q_offsets", ctypes.c_void_p),
        ("q", ctypes.c_void_p),
        ("form_factor", ctypes.c_void_p),
        ("atomic_number", ctypes.c_void_p),
        ("spectrum_energies", ctypes.c_void_p),
        ("spectrum_cdf", ctypes.c_void_p),
    ]


class _Source(ctypes.Structure):
    _fields_ = [
        ("source", ctypes.c_double * 3),
        ("corner", ctypes.c_double * 3),
        ("u", ctypes.c_double * 3),
        ("v", ctypes.c_double * 3),
        ("extent", ctypes.c_double * 2),
        ("weight", ctypes.c_double),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def _lib():
    global _LIB
    if _LIB is None:
        if not _LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(_LIB_PATH))
        vp, u64, u32, i64, dbl = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int64, ctypes.c_double
        lib.tr_random4.argtypes = [u64, u64, u32, u32, vp]
        lib.tr_random4.restype = None
        lib.tr_philox_words.argtypes = [u64, u64, u32, u32, vp]
        lib.tr_philox_words.restype = None
        lib.tr_trace.argtypes = [ctypes.POINTER(_Model), vp, ctypes.POINTER(_Source), u64, u64, i64, dbl,
                                 vp, vp, vp, vp, vp, vp, vp, vp]
        lib.tr_trace.restype = ctypes.c_int32
        lib.tr_scatter_direction.argtypes = [vp, dbl, dbl]
        lib.tr_scatter_direction.restype = None
        lib.tr_detector_pixel.argtypes = [ctypes.POINTER(_Model), vp, vp]
        lib.tr_detector_pixel.restype = ctypes.c_int32
        lib.tr_cell_index.argtypes = [ctypes.POINTER(_Model), vp, vp, vp]
        lib.tr_cell_index.restype = None
        lib.tr_compton_cosine.argtypes = [ctypes.POINTER(_Model), dbl, u64, u64, u32, vp]
        lib.tr_compton_cosine.restype = dbl
        lib.tr_rayleigh_cosine.argtypes = [ctypes.POINTER(_Model), dbl, ctypes.c_int, u64, u64, u32]
        lib.tr_rayleigh_cosine.restype = dbl
        lib.tr_threads.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def threads() -> int:
    return int(_lib().tr_threads())


def philox_words(seed: int, identity: int, event: int, domain: int) -> tuple[int, int, int, int]:
    out = np.zeros(4, dtype=np.uint32)
    _lib().tr_philox_words(seed, identity, event, domain, out.ctypes.data)
    return tuple(int(x) for x in out)


def random4(seed: int, identity: int, event: int, domain: int) -> np.ndarray:
    out = np.zeros(4, dtype=np.float64)
    _lib().tr_random4(seed, identity, event, domain, out.ctypes.data)
    return out


class Model:
    """ctypes view of a bound model; keeps its arrays alive."""

    def __init__(self, arrays: dict, scalars: dict) -> None:
        self._keep = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        m = _Model()
        m.nx, m.ny, m.nz = scalars["shape_xyz"]
        for key in ("materials", "energy_bins", "spectrum_bins", "max_events", "max_crossings", "max_angle_trials",
                    "compton", "compton_packed", "energy_score", "physics", "woodcock"):
            setattr(m, key, int(scalars[key]))
        for key in ("det_shape", "origin", "spacing", "det_corner", "det_u", "det_v", "det_normal", "det_spacing",
                    "inv_rot", "inv_trans"):
            target = getattr(m, key)
            for i, value in enumerate(scalars[key]):
                target[i] = value
        m.source_energy = float(scalars["source_energy"])
        m.cutoff = float(scalars["cutoff"])
        for key, value in self._keep.items():
            setattr(m, key, value.ctypes.data)
        self.struct = m
        self.pixels = int(scalars["det_shape"][0]) * int(scalars["det_shape"][1])


def _source(source) -> _Source:
    s = _Source()
    g = source.geometry
    for i in range(3):
        s.source[i] = g.source_mm[i]
        s.corner[i] = source.corner_mm[i]
        s.u[i] = g.u[i]
        s.v[i] = g.v[i]
    s.extent[0], s.extent[1] = source.extent_mm
    s.weight = float(source.weight)
    return s


def trace(model: Model, density, source, seed: int, first: int, count: int, *, amplitude: float = 1.0,
          records: bool = True, tally: bool = True) -> dict:
    """Walk histories first..first+count-1; returns per-history arrays (when
    ``records``), the (4, pixels) tally, the 7-count histogram and status bits."""
    dens = np.ascontiguousarray(density, dtype=np.float64)
    src = _source(source)
    out = {}
    ptr = {}
    if records:
        for key, dtype in (("pixel", np.int32), ("score", np.float64), ("energy", np.float64),
                           ("events", np.int32), ("real_events", np.int32), ("status", np.int32)):
            out[key] = np.zeros(count, dtype=dtype)
            ptr[key] = out[key].ctypes.data
    tally_array = np.zeros(4 * model.pixels, dtype=np.float64) if tally else None
    histogram = np.zeros(7, dtype=np.int64)
    bits = _lib().tr_trace(
        ctypes.byref(model.struct), dens.ctypes.data, ctypes.byref(src), seed, first, count, float(amplitude),
        ptr.get("pixel"), ptr.get("score"), ptr.get("energy"), ptr.get("events"), ptr.get("real_events"),
        ptr.get("status"), tally_array.ctypes.data if tally else None, histogram.ctypes.data,
    )
    if tally:
        out["tally"] = tally_array.reshape(4, model.pixels)
    out["histogram"] = histogram
    out["bits"] = int(bits)
    return out


def moments(tally: np.ndarray, histories: int) -> tuple[np.ndarray, np.ndarray]:
    """(3, pixels) mean and variance of the mean for total, primary, scatter
    (transport/simulation.py:61-66), from the (4, pixels) tally."""
    s = np.stack([tally[0] + tally[2], tally[0], tally[2]])
    q = np.stack([tally[1] + tally[3], tally[1], tally[3]])
    n = float(histories)
    mean = s / n
    m2 = np.maximum(q - s * mean, 0.0)
    return mean, m2 / (n * (n - 1.0))


def compton_cosine(model: Model, energy: float, seed: int, history: int, event: int) -> tuple[float, float]:
    out = ctypes.c_double(0.0)
    c = _lib().tr_compton_cosine(ctypes.byref(model.struct), energy, seed, history, event, ctypes.byref(out))
    return float(c), float(out.value)


def rayleigh_cosine(model: Model, energy: float, material: int, seed: int, history: int, event: int) -> float:
    return float(_lib().tr_rayleigh_cosine(ctypes.byref(model.struct), energy, material, seed, history, event))


def detector_pixel(model: Model, position, direction) -> int:
    p = np.ascontiguousarray(position, dtype=np.float64)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    return int(_lib().tr_detector_pixel(ctypes.byref(model.struct), p.ctypes.data, d.ctypes.data))


def cell_index(model: Model, position, direction) -> tuple[int, int, int]:
    p = np.ascontiguousarray(position, dtype=np.float64)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    out = np.zeros(3, dtype=np.int32)
    _lib().tr_cell_index(ctypes.byref(model.struct), p.ctypes.data, d.ctypes.data, out.ctypes.data)
    return tuple(int(x) for x in out)


def scatter_direction(direction, cosine: float, azimuth: float) -> np.ndarray:
    d = np.ascontiguousarray(direction, dtype=np.float64).copy()
    _lib().tr_scatter_direction(d.ctypes.data, cosine, azimuth)
    return d
