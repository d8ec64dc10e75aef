"""Continuous-batching pose-fit service on resident volumes.

The reference has no serving path: it runs one pose solve per process
(python/dpt/pose_optimization.py:753-1006).  This module keeps volumes and
per-solve state on the device, admits solves as they arrive when a slot and
HBM budget are free, and advances *all* active solves together: each step is
one ``dpt_fit_step`` launch (a Levenberg-Marquardt iteration of every active
solve; one ray march per ray gives the residual and the pose Jacobian row),
optionally replayed from a captured hipGraph.  A solve that finishes frees its
slot and memory at the next step; solves that arrive meanwhile join the batch.

Per-request results do not depend on the batch: each ray's arithmetic is
independent of the other requests and each request's normal equations are
summed in a fixed order, so a solve returns the same bits alone and in a mix.
"""

from __future__ import annotations

import ctypes
import math
import threading
import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import _native
from .contracts import ContractError
from .geometry import BatchedGeometry, RigidTransform
from .volumes import GridSpec

STATE_SIZE = 56
POSE, BEST_POSE, BEST_LOSS, LAMBDA, EVALS, ACCEPTED, LAST_LOSS = 0, 12, 24, 25, 26, 27, 55
STORAGE_DTYPES = {"float32": "float32", "bfloat16": "bfloat16", "float16": "float16"}


class VolumeDesc(ctypes.Structure):
    _fields_ = [
        ("field", ctypes.c_void_p),
        ("storage", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("object_to_grid", ctypes.c_double * 9),
        ("grid_origin", ctypes.c_double * 3),
    ]


class RequestDesc(ctypes.Structure):
    _fields_ = [
        ("volume", ctypes.c_int32),
        ("views", ctypes.c_int32),
        ("rows", ctypes.c_int32),
        ("cols", ctypes.c_int32),
        ("samples", ctypes.c_int32),
        ("panels", ctypes.c_void_p),
        ("targets", ctypes.c_void_p),
        ("state", ctypes.c_void_p),
    ]


def _bind(lib):
    if getattr(lib, "_dpt_fit_bound", False):
        return lib
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    lib.dpt_fit_create.argtypes = [i32, i32, i64, ctypes.POINTER(vp)]
    lib.dpt_fit_create.restype = ctypes.c_int
    lib.dpt_fit_destroy.argtypes = [vp]
    lib.dpt_fit_destroy.restype = ctypes.c_int
    lib.dpt_fit_step.argtypes = [vp, ctypes.POINTER(VolumeDesc), i32, ctypes.POINTER(RequestDesc), i32, i32, vp]
    lib.dpt_fit_step.restype = ctypes.c_int
    lib.dpt_fit_last_error.argtypes = []
    lib.dpt_fit_last_error.restype = ctypes.c_char_p
    lib._dpt_fit_bound = True
    return lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        message = _native.lib().dpt_fit_last_error().decode(errors="replace")
        if rc == -1:
            raise ContractError(f"{what}: {message}")
        raise _native.NativeLibraryError(f"{what} failed ({rc}): {message}")


@dataclass
class ResidentVolume:
    key: str
    grid: GridSpec
    storage: str
    tensor: object  # torch tensor on the device (flat)
    layout: str = "linear"  # "linear" (nz, ny, nx) or "cells" (dpt_convert_cells)

    @property
    def nbytes(self) -> int:
        return int(self.tensor.numel() * self.tensor.element_size())

    def desc(self) -> VolumeDesc:
        d = VolumeDesc()
        d.field = self.tensor.data_ptr()
        d.storage = _native.STORAGE_CODES[self.storage] | (_native.STORE_CELLS if self.layout == "cells" else 0)
        d.nx, d.ny, d.nz = self.grid.xyz
        for i, v in enumerate(self.grid.object_to_grid_matrix()):
            d.object_to_grid[i] = v
        for i in range(3):
            d.grid_origin[i] = self.grid.origin_mm[i]
        return d


class VolumeRegistry:
    """Device-resident volumes keyed by id, stored FP32, BF16 or FP16, in the
    linear (nz, ny, nx) layout or the cell layout (each interpolation cell's 8
    corners contiguous: one or two 16-byte loads per trilinear gather; 8x the
    linear footprint, built once when the volume is added)."""

    def __init__(self, device="cuda:0") -> None:
        import torch

        self.device = torch.device(device)
        self._volumes: dict[str, ResidentVolume] = {}
        self._order: list[str] = []

    def add(self, key: str, values, grid: GridSpec, storage: str = "bfloat16", layout: str = "cells") -> ResidentVolume:
        import torch

        if storage not in STORAGE_DTYPES:
            raise ContractError("storage must be float32, bfloat16 or float16")
        if layout not in ("linear", "cells"):
            raise ContractError("layout must be linear or cells")
        src = torch.as_tensor(np.ascontiguousarray(values, dtype=np.float32)).to(self.device).reshape(-1)
        if tuple(np.shape(values)) != grid.shape:
            raise ContractError("values must have the grid shape")
        if not bool(torch.isfinite(src).all()) or float(src.min()) < 0.0:
            raise ContractError("resident attenuation must be finite and non-negative")
        if storage == "float32":
            tensor = src
        else:
            tensor = torch.empty(src.numel() + _native.HALF_PAD_ELEMENTS, dtype=getattr(torch, storage), device=self.device)
            stream = torch.cuda.current_stream(self.device).cuda_stream
            _native.convert_field(src.data_ptr(), tensor.data_ptr(), src.numel(), storage, None, stream)
        if layout == "cells":
            cells = torch.empty(8 * _native.cell_count(grid.shape), dtype=tensor.dtype, device=self.device)
            stream = torch.cuda.current_stream(self.device).cuda_stream
            _native.convert_cells(tensor.data_ptr(), cells.data_ptr(), grid.shape, storage, stream)
            tensor = cells
        # resident data must be complete before any engine stream reads it
        torch.cuda.current_stream(self.device).synchronize()
        volume = ResidentVolume(key, grid, storage, tensor, layout)
        if key not in self._volumes:
            self._order.append(key)
        self._volumes[key] = volume
        return volume

    def __getitem__(self, key: str) -> ResidentVolume:
        return self._volumes[key]

    def __contains__(self, key: str) -> bool:
        return key in self._volumes

    @property
    def keys(self) -> list[str]:
        return list(self._order)

    @property
    def nbytes(self) -> int:
        return sum(v.nbytes for v in self._volumes.values())


@dataclass
class FitRequest:
    """One pose-fit solve: fit the object pose so projections match targets."""

    key: str
    volume: str
    geometry: BatchedGeometry
    targets: np.ndarray  # (views, rows, cols) measured optical depth
    samples: int
    iterations: int
    initial_pose: tuple = RigidTransform().packed()
    arrival: float = 0.0  # seconds after the run starts

    @property
    def views(self) -> int:
        return len(self.geometry.views)

    @property
    def nbytes(self) -> int:
        v, r, c = self.geometry.shape
        return v * r * c * 4 + v * 12 * 8 + STATE_SIZE * 8

    @property
    def tiles(self) -> int:
        v, r, c = self.geometry.shape
        return v * math.ceil(r / 16) * math.ceil(c / 16)


@dataclass
class FitResult:
    key: str
    pose: np.ndarray  # (12,) best packed pose
    loss: float
    evaluations: int
    accepted: int
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Active:
    request: FitRequest
    targets: object
    panels: object
    state: object
    done: int = 0
    admitted: float = 0.0
    staging: tuple = ()  # pinned host sources of the asynchronous uploads
    serial: int = 0      # unique per admission (descriptor cache key)


@dataclass
class RunStats:
    results: list[FitResult]
    makespan: float
    steps: int
    peak_active: int
    max_bytes: int

    @property
    def solves_per_s(self) -> float:
        return len(self.results) / self.makespan if self.makespan > 0 else 0.0

    def latency_percentile(self, q: float) -> float:
        values = np.asarray([r.latency for r in self.results])
        return float(np.percentile(values, q)) if values.size else math.nan

    def summary(self) -> dict:
        return {
            "solves": len(self.results),
            "makespan_s": round(self.makespan, 4),
            "solves_per_s": round(self.solves_per_s, 3),
            "p50_latency_ms": round(self.latency_percentile(50) * 1e3, 3),
            "p99_latency_ms": round(self.latency_percentile(99) * 1e3, 3),
            "steps": self.steps,
            "peak_active": self.peak_active,
            "peak_request_bytes": self.max_bytes,
        }


class PoseFitEngine:
    """Admit, batch and complete pose-fit solves on one GPU.

    ``hbm_budget`` bounds the bytes of per-request device state (targets,
    panels, solver state); ``max_requests`` bounds the batch.  ``use_graph``
    replays each step from a hipGraph captured per grid size.
    """

    def __init__(
        self,
        registry: VolumeRegistry,
        *,
        hbm_budget: int = 8 << 30,
        max_requests: int = 64,
        max_tiles: int = 1 << 19,
        use_graph: bool = True,
    ) -> None:
        import torch

        self.registry = registry
        self.device = registry.device
        self.hbm_budget = int(hbm_budget)
        self.max_requests = int(max_requests)
        self.max_tiles = int(max_tiles)
        self.use_graph = bool(use_graph)
        self._lib = _bind(_native.lib())
        handle = ctypes.c_void_p()
        _check(
            self._lib.dpt_fit_create(self.device.index or 0, self.max_requests, self.max_tiles, ctypes.byref(handle)),
            "dpt_fit_create",
        )
        self._handle = handle
        # graph capture needs a non-default stream (default priority)
        self.stream = torch.cuda.Stream(device=self.device)
        self._desc_cache = ((), None)
        self._serial = 0
        self._lock = threading.Lock()

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None and self._handle.value:
            self._lib.dpt_fit_destroy(self._handle)
            self._handle = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    def _admit(self, request: FitRequest, now: float) -> _Active:
        import torch

        if request.volume not in self.registry:
            raise ContractError(f"request {request.key}: volume {request.volume!r} is not resident")
        v, r, c = request.geometry.shape
        # pinned host copies uploaded asynchronously on the engine stream: an
        # admission does not wait for the steps already in flight
        host_targets = torch.as_tensor(np.ascontiguousarray(request.targets, dtype=np.float32)).reshape(v, r, c)
        host_targets = host_targets.pin_memory()
        host_panels = torch.tensor([p.packed() for p in request.geometry.views], dtype=torch.float64).pin_memory()
        host_state = torch.zeros(STATE_SIZE, dtype=torch.float64)
        host_state[POSE : POSE + 12] = torch.tensor(request.initial_pose, dtype=torch.float64)
        host_state = host_state.pin_memory()
        targets = host_targets.to(self.device, non_blocking=True)
        panels = host_panels.to(self.device, non_blocking=True)
        state = host_state.to(self.device, non_blocking=True)
        self._serial += 1
        return _Active(request, targets, panels, state, 0, now, (host_targets, host_panels, host_state), self._serial)

    def _descs(self, active: list[_Active]):
        signature = tuple(a.serial for a in active)
        if signature == self._desc_cache[0]:
            return self._desc_cache[1]
        keys = sorted({a.request.volume for a in active})
        index = {k: i for i, k in enumerate(keys)}
        vols = (VolumeDesc * len(keys))(*[self.registry[k].desc() for k in keys])
        reqs = (RequestDesc * len(active))()
        for i, a in enumerate(active):
            q = reqs[i]
            q.volume = index[a.request.volume]
            q.views, q.rows, q.cols = a.request.geometry.shape
            q.samples = a.request.samples
            q.panels = a.panels.data_ptr()
            q.targets = a.targets.data_ptr()
            q.state = a.state.data_ptr()
        self._desc_cache = (signature, (vols, len(keys), reqs))
        return vols, len(keys), reqs

    def step(self, active: list[_Active]) -> None:
        """One Levenberg-Marquardt iteration of every active solve (async)."""
        if not active:
            return
        # longest rays first: the dynamic tile counter then balances the tail
        order = sorted(active, key=lambda a: -a.request.samples)
        vols, nv, reqs = self._descs(order)
        _check(
            self._lib.dpt_fit_step(
                self._handle, vols, nv, reqs, len(order), 1 if self.use_graph else 0, self.stream.cuda_stream
            ),
            "dpt_fit_step",
        )

    def run(self, requests: list[FitRequest], *, realtime: bool = True, progress=None) -> RunStats:
        """Serve requests arriving at ``request.arrival`` seconds (continuous batching).

        With ``realtime`` the run waits for arrivals on the wall clock; without
        it every request is queued at t = 0 (closed-loop throughput).
        ``progress``, if given, is called with the number of completed
        requests each time solves complete.
        """
        import torch

        pending = deque(sorted(requests, key=lambda q: q.arrival))
        queued: deque[FitRequest] = deque()
        active: list[_Active] = []
        results: list[FitResult] = []
        used = 0
        steps = peak = peak_bytes = 0
        inflight: deque = deque()
        completions: deque = deque()
        start = time.perf_counter()

        def emit(landed, snapshot, finished):
            stop = time.perf_counter() - start
            states = snapshot.numpy()
            for a, st in zip(finished, states):
                results.append(
                    FitResult(
                        a.request.key,
                        st[BEST_POSE : BEST_POSE + 12].copy(),
                        float(st[BEST_LOSS]),
                        int(st[EVALS]),
                        int(st[ACCEPTED]),
                        a.request.arrival if realtime else 0.0,
                        a.admitted,
                        stop,
                    )
                )
            if progress is not None and finished:
                progress(len(results))

        with self._lock, torch.cuda.stream(self.stream):
            while pending or queued or active:
                while completions and completions[0][0].query():
                    emit(*completions.popleft())
                now = time.perf_counter() - start
                while pending and (not realtime or pending[0].arrival <= now):
                    queued.append(pending.popleft())
                # admission: slot, tile capacity and HBM budget
                tiles = sum(a.request.tiles for a in active)
                while queued and len(active) < self.max_requests:
                    q = queued[0]
                    if used + q.nbytes > self.hbm_budget and active:
                        break
                    if tiles + q.tiles > self.max_tiles and active:
                        break
                    queued.popleft()
                    active.append(self._admit(q, now))
                    used += q.nbytes
                    tiles += q.tiles
                peak = max(peak, len(active))
                peak_bytes = max(peak_bytes, used)
                if not active:
                    if pending:
                        time.sleep(max(0.0, min(pending[0].arrival - now, 0.002)))
                    continue
                self.step(active)
                steps += 1
                for a in active:
                    a.done += 1
                finished = [a for a in active if a.done >= a.request.iterations]
                if finished:
                    # The final states are copied to pinned memory behind the
                    # step on the same stream; the solves leave the batch now
                    # and their results are emitted when the copy has landed,
                    # without draining the steps in flight.
                    snapshot = torch.empty((len(finished), STATE_SIZE), dtype=torch.float64, pin_memory=True)
                    snapshot.copy_(torch.stack([a.state for a in finished]), non_blocking=True)
                    landed = torch.cuda.Event()
                    landed.record(self.stream)
                    completions.append((landed, snapshot, finished))
                    for a in finished:
                        used -= a.request.nbytes
                    active = [a for a in active if a.done < a.request.iterations]
                # keep at most two steps in flight (the native table staging is
                # double-buffered)
                event = torch.cuda.Event()
                event.record(self.stream)
                inflight.append(event)
                if len(inflight) > 2:
                    inflight.popleft().synchronize()
            while completions:
                completions[0][0].synchronize()
                emit(*completions.popleft())
        return RunStats(results, time.perf_counter() - start, steps, peak, peak_bytes)


# ------------------------------------------------------------- workloads ----


def poisson_arrivals(count: int, rate: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


def rotation_angles(xi) -> tuple[float, float]:
    """(rotation angle in degrees, translation norm in mm) of a 6-vector."""
    return math.degrees(math.sqrt(sum(x * x for x in xi[3:]))), math.sqrt(sum(x * x for x in xi[:3]))


def pose_error(a, b) -> tuple[float, float]:
    """Rotation (degrees) and translation (mm) distance between packed poses."""
    ra = np.asarray(a[:9]).reshape(3, 3)
    rb = np.asarray(b[:9]).reshape(3, 3)
    c = (np.trace(ra.T @ rb) - 1.0) / 2.0
    angle = math.degrees(math.acos(max(-1.0, min(1.0, c))))
    return angle, float(np.linalg.norm(np.asarray(a[9:]) - np.asarray(b[9:])))
