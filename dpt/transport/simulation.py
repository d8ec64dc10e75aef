"""Prepared transport, checked simulation and per-history traces on the GPU.

Host side of the reference's transport path: prepare_transport
(python/dpt/transport/forward.py:146-264), prepare_simulation / simulate
(transport/simulation.py:121-316) and trace_histories
(transport/forward.py:330-425).  The walk itself runs in the HIP kernel
behind dpt_transport_trace; this module validates records, binds tables,
owns device buffers (torch tensors as plumbing) and maps statuses to errors
exactly as TransportWorkspace._raise_status does (transport/forward.py:93-103).
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from .. import _native
from ..contracts import NumericalError
from .model import (
    BUDGET_MASK,
    ENERGY_MASK,
    NUMERICAL_MASK,
    HistoryStatus,
    IncompleteHistoryError,
    OrientedDetector,
    TransportError,
    TransportSpec,
)
from .physical import HC_KEV_ANGSTROM, PhysicalTransport
from .source import HistoryBatch, PointConeSource


def _torch():
    import torch

    return torch


def _as_numpy(value, dtype, name: str) -> np.ndarray:
    if hasattr(value, "detach"):
        value = value.detach().cpu().numpy()
    array = np.ascontiguousarray(np.asarray(value), dtype=dtype).reshape(-1)
    return array


def _check_transport(rc: int, what: str) -> None:
    """Native status -> TransportError for contract violations (DPT_E_INVALID)."""
    if rc == -1:
        message = _native.lib().dpt_transport_last_error().decode(errors="replace")
        raise TransportError(f"{what}: {message}")
    _native.check(rc, what, _native.lib().dpt_transport_last_error)


def bind_model(spec: TransportSpec, *, material_ids, absorption, scattering,
               physical: PhysicalTransport | None = None, max_histories: int = 1) -> tuple[dict, dict]:
    """Validate the records and flatten them into the bound model: host arrays
    (labels as 8-bit, tables material-major with energy fastest) and scalars
    in the object frame, exactly what dpt_transport_create receives."""
    if not isinstance(spec, TransportSpec):
        raise TransportError("spec must be a TransportSpec")
    if type(max_histories) is not int or not 0 < max_histories < 2**31:
        raise TransportError("max_histories must be a positive signed 32-bit integer")
    grid = spec.grid
    voxels = math.prod(grid.shape)
    ids = _as_numpy(material_ids, np.int64, "material_ids")
    if ids.size != voxels:
        raise TransportError(f"material_ids must have {voxels} entries (x fastest)")
    if ids.min() < 0 or ids.max() >= grid.materials:
        raise TransportError("material ids must lie in 0..materials-1")
    if grid.materials > _DEVICE_MATERIALS:
        raise TransportError(f"at most {_DEVICE_MATERIALS} materials are supported")
    nodes = np.asarray(spec.energy_nodes, dtype=np.float64)
    K, M = nodes.size, grid.materials
    a = _as_numpy(absorption, np.float64, "absorption")
    s = _as_numpy(scattering, np.float64, "scattering")
    for name, table in (("absorption", a), ("scattering", s)):
        if table.size != M * K:
            raise TransportError(f"{name} must have materials*energies = {M * K} entries")
        if not np.all(np.isfinite(table)) or np.any(table < 0):
            raise TransportError(f"{name} coefficients must be finite and nonnegative")
    present = np.zeros(M, dtype=np.float64)
    present[np.unique(ids)] = 1.0
    arrays = {
        "material_ids": ids.astype(np.uint8),
        "energies": nodes,
        "absorption": a,
        "scattering": s,
        "majorant_max_density": present,
    }
    physics = physical is not None
    woodcock = False
    spectrum_bins = 0
    if physics:
        if not isinstance(physical, PhysicalTransport):
            raise TransportError("physical must be a PhysicalTransport record")
        if len(physical.scattering_tables) != M:
            raise TransportError("physical transport needs one scattering table per material")
        if len(physical.rayleigh_mm_inverse) != M * K:
            raise TransportError("Rayleigh table must match the material/energy coefficient shape")
        if physical.majorant_mm_inverse and len(physical.majorant_mm_inverse) != K:
            raise TransportError("Woodcock majorant must match the coefficient energy nodes")
        if physical.cutoff_kev > nodes[-1]:
            raise TransportError("cutoff exceeds the supplied energy support")
        if spec.max_angle_trials > 65536:
            raise TransportError("physical angular rejection reserves at most 65536 trial domains")
        spectrum = physical.spectrum.energy_kev if physical.spectrum else ()
        if spectrum and (spectrum[0] < nodes[0] or spectrum[-1] > nodes[-1]):
            raise TransportError("source spectrum leaves coefficient energy support")
        e_max = max((spec.energy_kev, *spectrum))
        rayleigh = np.asarray(physical.rayleigh_mm_inverse, dtype=np.float64)
        if np.any(rayleigh > s):
            raise TransportError("the Rayleigh partial exceeds aggregate scattering")
        offsets, q, ff, z = [0], [], [], []
        for table in physical.scattering_tables:
            if table.momentum_transfer_inverse_angstrom[-1] < e_max / HC_KEV_ANGSTROM:
                raise TransportError("atomic scattering q support does not cover all source angles")
            q.extend(table.momentum_transfer_inverse_angstrom)
            ff.extend(table.form_factor)
            z.append(float(table.atomic_number))
            offsets.append(len(q))
        arrays.update(
            rayleigh=rayleigh,
            q_offsets=np.asarray(offsets, dtype=np.int32),
            q=np.asarray(q, dtype=np.float64),
            form_factor=np.asarray(ff, dtype=np.float64),
            atomic_number=np.asarray(z, dtype=np.float64),
        )
        if physical.majorant_mm_inverse:
            woodcock = True
            arrays["majorant"] = np.asarray(physical.majorant_mm_inverse, dtype=np.float64)
        if physical.spectrum is not None:
            spectrum_bins = len(physical.spectrum.energy_kev)
            arrays["spectrum_energies"] = np.asarray(physical.spectrum.energy_kev, dtype=np.float64)
            arrays["spectrum_cdf"] = np.asarray(physical.spectrum.cdf, dtype=np.float64)
    inverse = spec.object_pose.inverse()
    det = spec.detector
    scalars = {
        "shape_xyz": grid.shape_xyz,
        "materials": M,
        "energy_bins": K,
        "spectrum_bins": spectrum_bins,
        "max_events": spec.max_events,
        "max_crossings": spec.max_crossings,
        "max_angle_trials": spec.max_angle_trials,
        "compton": int(spec.scattering_law == "free-electron-compton"),
        "compton_packed": int(spec.compton_sampler == "packed-uniform-rejection"),
        "energy_score": int(spec.scoring == "energy-kev"),
        "physics": int(physics),
        "woodcock": int(woodcock),
        "det_shape": tuple(det.shape),
        "origin": tuple(grid.origin_mm),
        "spacing": tuple(grid.spacing_mm),
        "det_corner": inverse.point(det.lower_corner_mm),
        "det_u": inverse.direction(det.u),
        "det_v": inverse.direction(det.v),
        "det_normal": inverse.direction(det.normal),
        "det_spacing": tuple(det.spacing_xy_mm),
        "inv_rot": tuple(inverse.rotation),
        "inv_trans": tuple(inverse.translation_mm),
        "source_energy": float(spec.energy_kev),
        "cutoff": float(physical.cutoff_kev) if physics else 0.0,
    }
    return arrays, scalars


class TransportWorkspace:
    """Bound model: labels and tables resident on one device (forward.py:44-143).

    ``arrays`` holds the host copies of every bound table (material-major,
    energy fastest) so that a checker can run the same model elsewhere.
    """

    def __init__(self, spec: TransportSpec, *, material_ids, absorption, scattering, max_histories: int,
                 device: int = 0, physical: PhysicalTransport | None = None) -> None:
        arrays, scalars = bind_model(spec, material_ids=material_ids, absorption=absorption, scattering=scattering,
                                     physical=physical, max_histories=max_histories)
        grid = spec.grid
        self.spec, self.physical, self.max_histories = spec, physical, max_histories
        self.arrays, self.scalars = arrays, scalars
        self.device = int(device)
        model = _native.TransportModel()
        model.nx, model.ny, model.nz = grid.shape_xyz
        for key in ("materials", "energy_bins", "spectrum_bins", "max_events", "max_crossings", "max_angle_trials",
                    "compton", "compton_packed", "energy_score", "physics", "woodcock"):
            setattr(model, key, scalars[key])
        for key in ("det_shape", "origin", "spacing", "det_corner", "det_u", "det_v", "det_normal", "det_spacing",
                    "inv_rot", "inv_trans"):
            target = getattr(model, key)
            for i, value in enumerate(scalars[key]):
                target[i] = value
        model.source_energy = scalars["source_energy"]
        model.cutoff = scalars["cutoff"]
        for key, value in arrays.items():
            setattr(model, key, value.ctypes.data)
        handle = ctypes.c_void_p()
        _check_transport(_native.lib().dpt_transport_create(ctypes.byref(model), self.device, ctypes.byref(handle)),
                         "prepare_transport")
        self.handle = handle

    @property
    def pixels(self) -> int:
        return self.spec.detector.pixels

    def _source(self, source: PointConeSource) -> _native.PointSource:
        if not isinstance(source, PointConeSource):
            raise TransportError("simulate requires a library PointConeSource")
        if self.spec.detector.source_distance(source.geometry.source_mm) >= 0:
            raise TransportError("source batch capacity or source-detector geometry is invalid")
        s = _native.PointSource()
        for i in range(3):
            s.source[i] = source.geometry.source_mm[i]
            s.corner[i] = source.corner_mm[i]
            s.u[i] = source.geometry.u[i]
            s.v[i] = source.geometry.v[i]
        s.extent[0], s.extent[1] = source.extent_mm
        s.weight = float(source.weight)
        return s

    def _density(self, density) -> np.ndarray:
        values = _as_numpy(density, np.float64, "density")
        if values.size != self.spec.grid.materials:
            raise TransportError("density must have one positive scale per material")
        return values

    def trace(self, density, source: PointConeSource, batch: HistoryBatch, *, amplitude: float = 1.0,
              records: dict | None = None, tally=None, stream=None) -> None:
        """Enqueue the walk of every history in ``batch`` (no synchronisation)."""
        if not isinstance(batch, HistoryBatch):
            raise TransportError("simulate batch must be a HistoryBatch")
        if batch.count > self.max_histories:
            raise TransportError("history batch exceeds workspace capacity")
        torch = _torch()
        values = self._density(density)
        src = self._source(source)
        rec = None
        if records is not None:
            rec = _native.HistoryRecords()
            for key, tensor in records.items():
                setattr(rec, key, tensor.data_ptr())
        if stream is None:
            stream = torch.cuda.current_stream(self.device).cuda_stream
        _check_transport(
            _native.lib().dpt_transport_trace(
                self.handle, values.ctypes.data, ctypes.byref(src), batch.seed, batch.first_history, batch.count,
                float(amplitude), ctypes.byref(rec) if rec is not None else None,
                tally.data_ptr() if tally is not None else None, stream,
            ),
            "trace_histories",
        )

    def consume_status(self, stream=None) -> tuple[int, ...]:
        """Synchronise, read and clear the terminal counts; raise on failures."""
        torch = _torch()
        if stream is None:
            stream = torch.cuda.current_stream(self.device).cuda_stream
        counts = np.zeros(7, dtype=np.int64)
        bits = ctypes.c_int32(0)
        _check_transport(
            _native.lib().dpt_transport_status(self.handle, counts.ctypes.data, ctypes.byref(bits), 1, stream),
            "transport status",
        )
        code = int(bits.value)
        if code & BUDGET_MASK:
            raise IncompleteHistoryError(
                "event, voxel-crossing or angular-rejection budget exhausted; discard the estimate"
            )
        if code & ENERGY_MASK:
            raise TransportError("a history left the supplied coefficient energy support")
        if code & NUMERICAL_MASK:
            raise NumericalError("nonfinite transport arithmetic or an invalid moment")
        return tuple(int(c) for c in counts)

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            _native.lib().dpt_transport_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


_DEVICE_MATERIALS = 8


def prepare_transport(spec: TransportSpec, *, material_ids, absorption, scattering, max_histories: int,
                      device: int = 0, physical: PhysicalTransport | None = None) -> TransportWorkspace:
    """Bind labels and coefficient tables on one device (transport/forward.py:146-264)."""
    return TransportWorkspace(spec, material_ids=material_ids, absorption=absorption, scattering=scattering,
                              max_histories=max_histories, device=device, physical=physical)


@dataclass(frozen=True)
class HistoryStatusHistogram:
    """Seven exact counts ordered by HistoryStatus (transport/simulation.py:27-51)."""

    counts: tuple[int, int, int, int, int, int, int]

    @property
    def total(self) -> int:
        return sum(self.counts)

    def count(self, status: HistoryStatus) -> int:
        return self.counts[int(status)]


@dataclass(frozen=True)
class SimulationResult:
    """Per-pixel mean and variance of the mean (transport/simulation.py:54-90).

    Each image is a flat FP64 device tensor of ``pixels`` entries aliasing the
    workspace (the next simulation overwrites it).  ``mean`` /
    ``estimated_variance_of_mean`` score every detected history; the primary
    and scatter images split them by zero / nonzero real interactions
    (transport/readout.py:19-24).
    """

    mean: object
    estimated_variance_of_mean: object
    primary_mean: object
    primary_variance_of_mean: object
    scatter_mean: object
    scatter_variance_of_mean: object
    status_histogram: HistoryStatusHistogram
    batch: HistoryBatch
    image_shape: tuple[int, int]
    device: str
    estimator: str
    scoring: str


class SimulationWorkspace:
    """Tally, moment and image buffers for ``simulate`` (transport/simulation.py:93-183)."""

    def __init__(self, transport: TransportWorkspace, max_histories: int | None = None) -> None:
        if not isinstance(transport, TransportWorkspace):
            raise TransportError("prepare_simulation transport must be a TransportWorkspace")
        capacity = transport.max_histories if max_histories is None else max_histories
        if type(capacity) is not int or not 2 <= capacity <= transport.max_histories:
            raise TransportError("simulation capacity must be 2..transport.max_histories")
        torch = _torch()
        device = f"cuda:{transport.device}"
        pixels = transport.pixels
        self.transport = transport
        self.max_histories = capacity
        self.tally = torch.zeros(4 * pixels, dtype=torch.float64, device=device)
        self.mean = torch.empty(3 * pixels, dtype=torch.float64, device=device)
        self.variance = torch.empty(3 * pixels, dtype=torch.float64, device=device)


def prepare_simulation(transport: TransportWorkspace, *, max_histories: int | None = None) -> SimulationWorkspace:
    return SimulationWorkspace(transport, max_histories)


def simulate(density, *, source: PointConeSource, batch: HistoryBatch, workspace: SimulationWorkspace,
             source_amplitude: float = 1.0, chunk: int | None = None) -> SimulationResult:
    """Mean[p] = sum_i X_i[p] / N over original histories (transport/simulation.py:186-316).

    All histories, including misses and absorption, stay in N.  ``chunk``
    splits the batch into consecutive identity ranges traced one after another
    (results do not depend on it: the walk of a history depends only on its
    identity).  One synchronisation at the end reads the status histogram.
    """
    if not isinstance(workspace, SimulationWorkspace):
        raise TransportError("simulate workspace must be a SimulationWorkspace")
    if not isinstance(batch, HistoryBatch):
        raise TransportError("simulate batch must be a HistoryBatch")
    if not 2 <= batch.count <= workspace.max_histories:
        raise TransportError("simulation batch must contain 2..workspace.max_histories histories")
    if not math.isfinite(source_amplitude) or source_amplitude < 0:
        raise TransportError("source amplitude must be finite and nonnegative")
    torch = _torch()
    transport = workspace.transport
    stream = torch.cuda.current_stream(transport.device).cuda_stream
    workspace.tally.zero_()
    step = batch.count if chunk is None else max(1, int(chunk))
    for start in range(0, batch.count, step):
        part = HistoryBatch(batch.seed, batch.first_history + start, min(step, batch.count - start))
        transport.trace(density, source, part, amplitude=source_amplitude, tally=workspace.tally, stream=stream)
    _check_transport(
        _native.lib().dpt_transport_moments(
            transport.handle, workspace.tally.data_ptr(), batch.count, workspace.mean.data_ptr(),
            workspace.variance.data_ptr(), stream,
        ),
        "history moments",
    )
    counts = transport.consume_status(stream)
    histogram = HistoryStatusHistogram(counts)
    if histogram.total != batch.count:
        raise TransportError("terminal status histogram does not account for every history")
    p = transport.pixels
    m, v = workspace.mean, workspace.variance
    return SimulationResult(
        m[:p], v[:p], m[p:2 * p], v[p:2 * p], m[2 * p:], v[2 * p:], histogram, batch,
        tuple(transport.spec.detector.image_shape), f"cuda:{transport.device}", transport.spec.estimator,
        transport.spec.scoring,
    )


def trace_histories(density, *, source: PointConeSource, batch: HistoryBatch, workspace: TransportWorkspace,
                    source_amplitude: float = 1.0) -> dict:
    """Per-history pixel, score, energy, events, real events and status
    (transport/forward.py:330-425; kernels.py:1944-1994) as device tensors."""
    torch = _torch()
    device = f"cuda:{workspace.device}"
    n = batch.count
    records = {
        "pixel": torch.empty(n, dtype=torch.int32, device=device),
        "score": torch.empty(n, dtype=torch.float64, device=device),
        "energy": torch.empty(n, dtype=torch.float64, device=device),
        "events": torch.empty(n, dtype=torch.int32, device=device),
        "real_events": torch.empty(n, dtype=torch.int32, device=device),
        "status": torch.empty(n, dtype=torch.int32, device=device),
    }
    workspace.trace(density, source, batch, amplitude=source_amplitude, records=records)
    counts = workspace.consume_status()
    records["histogram"] = counts
    return records
