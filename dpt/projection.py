"""Host operator layer: configuration encoding, the plan cache and checkpoints.

Replaces the reference's per-view workspace stack
(python/dpt/projection.py:212-624 and python/dpt/torch/_projection_ops.py:
35-175): one native multiview plan per (configuration, device, stream, active
gradients), cached LRU-8 and guarded by a per-entry lock; one status readback
per composed operation (python/dpt/_runtime.py:317-445).
"""

from __future__ import annotations

import json
import math
import os
import threading
from collections import OrderedDict
from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

from . import _native
from .contracts import ContractError, integer, raise_view_statuses
from .geometry import BatchedGeometry, CurvedDetectorGeometry, DetectorGeometry, ParallelGeometry, RayBundleGeometry
from .volumes import GridSpec

CACHE_LIMIT = 8
PRECISIONS = ("float32", "float64")
INTEGRATIONS = ("midpoint", "cell_gauss")
STORAGES = ("float32", "bfloat16", "float16")


@dataclass(frozen=True)
class ProjectionSpec:
    """Midpoint samples, sample arithmetic, resident field storage and the
    integration rule.

    ``precision`` is the sampling and output arithmetic; ``storage`` is how the
    field is held on the device.  Half storage pairs only with FP32 sampling
    and uses compensated along-ray sums.  ``integration`` "cell_gauss" is the
    reference's exact rule (two Gauss nodes per interpolation cell,
    python/dpt/projection.py:107-135); it requires FP64 sampling and ignores
    ``samples_per_ray``.
    """

    samples_per_ray: int
    precision: str = "float32"
    storage: str = "float32"
    integration: str = "midpoint"

    def __post_init__(self) -> None:
        integer(self.samples_per_ray, "samples_per_ray", minimum=1)
        if self.precision not in PRECISIONS:
            raise ContractError("projection precision must be float32 or float64")
        if self.storage not in STORAGES:
            raise ContractError("storage must be float32, bfloat16 or float16")
        if self.storage != "float32" and self.precision != "float32":
            raise ContractError("half storage requires FP32 sampling")
        if self.integration not in INTEGRATIONS:
            raise ContractError("projection integration must be midpoint or cell_gauss")
        if self.integration == "cell_gauss" and self.precision != "float64":
            raise ContractError("cell_gauss integration requires float64 sampling")


def _encode_panel(p) -> dict:
    common = {"origin_mm": list(p.origin_mm), "u": list(p.u), "v": list(p.v), "spacing_mm": list(p.spacing_mm),
              "shape": list(p.shape)}
    if isinstance(p, ParallelGeometry):
        return {"kind": "parallel", "direction": list(p.direction), "length_mm": p.length_mm, **common}
    kind = "curved" if isinstance(p, CurvedDetectorGeometry) else "cone"
    return {"kind": kind, "source_mm": list(p.source_mm), **common}


def _decode_panel(p: dict):
    common = (tuple(p["u"]), tuple(p["v"]), tuple(p["spacing_mm"]), tuple(p["shape"]))
    kind = p.get("kind", "cone")
    if kind == "parallel":
        return ParallelGeometry(tuple(p["origin_mm"]), tuple(p["direction"]), p["length_mm"], *common)
    cls = CurvedDetectorGeometry if kind == "curved" else DetectorGeometry
    return cls(tuple(p["source_mm"]), tuple(p["origin_mm"]), *common)


def configuration(grid: GridSpec, geometry, spec: ProjectionSpec) -> str:
    """Encode fixed physical values as a JSON string (dispatcher-safe)."""
    if isinstance(geometry, RayBundleGeometry):
        views = {"rays": [list(map(list, geometry.sources_mm)), list(map(list, geometry.endpoints_mm))]}
    else:
        views = {"panels": [_encode_panel(p) for p in geometry.views]}
    return json.dumps(
        {
            "grid": grid.encode(),
            **views,
            "samples": spec.samples_per_ray,
            "precision": spec.precision,
            "storage": spec.storage,
            "integration": spec.integration,
        },
        sort_keys=True,
        separators=(",", ":"),
        allow_nan=False,
    )


@dataclass(frozen=True)
class Decoded:
    grid: GridSpec
    geometry: object  # BatchedGeometry | RayBundleGeometry
    spec: ProjectionSpec

    @property
    def image_shape(self) -> tuple[int, int, int]:
        if isinstance(self.geometry, RayBundleGeometry):
            return (1, 1, self.geometry.rays)
        return self.geometry.shape

    @property
    def views(self) -> int:
        return self.image_shape[0]


@lru_cache(maxsize=32)
def decode(value: str) -> Decoded:
    encoded = json.loads(value)
    grid = GridSpec.decode(encoded["grid"])
    if "rays" in encoded:
        sources, endpoints = encoded["rays"]
        geometry = RayBundleGeometry(tuple(map(tuple, sources)), tuple(map(tuple, endpoints)))
    else:
        geometry = BatchedGeometry(tuple(_decode_panel(p) for p in encoded["panels"]))
    spec = ProjectionSpec(int(encoded["samples"]), encoded["precision"], encoded.get("storage", "float32"),
                          encoded.get("integration", "midpoint"))
    return Decoded(grid, geometry, spec)


@dataclass
class Prepared:
    plan: _native.Plan
    decoded: Decoded
    lock: threading.RLock = field(default_factory=threading.RLock)
    half: object = None  # cached half-storage device buffer (torch tensor)


_CACHE: OrderedDict[tuple, Prepared] = OrderedDict()
_CACHE_LOCK = threading.RLock()


def _native_geometry(decoded: Decoded):
    geometry = decoded.geometry
    if isinstance(geometry, RayBundleGeometry):
        src = np.asarray(geometry.sources_mm, dtype=np.float64)
        end = np.asarray(geometry.endpoints_mm, dtype=np.float64)
        return dict(views=1, rows=1, cols=geometry.rays, panels=None, rays=(src, end))
    rows, cols = geometry.shape_rows_columns
    if not geometry.flat_cone:
        # parallel / cylindrical panels: one world segment per pixel, expanded on the device
        records = [p.record() for p in geometry.views]
        return dict(views=len(geometry.views), rows=rows, cols=cols, panels=None, rays=None,
                    panel_records=np.asarray([r for r, _ in records], dtype=np.float64),
                    panel_modes=np.asarray([m for _, m in records], dtype=np.int32))
    panels = np.asarray([p.packed() for p in geometry.views], dtype=np.float64)
    return dict(views=len(geometry.views), rows=rows, cols=cols, panels=panels, rays=None)


def _check_extent(decoded: Decoded) -> None:
    """Detector corners must give finite rays (projection.py:578-586)."""
    geometry = decoded.geometry
    if isinstance(geometry, RayBundleGeometry):
        return
    for view, panel in enumerate(geometry.views):
        for row in (0, panel.shape[0] - 1):
            for col in (0, panel.shape[1] - 1):
                source, end = geometry.segment(view, row, col)
                if not all(math.isfinite(x) for x in (*source, *end)) or not math.isfinite(math.dist(end, source)):
                    raise ContractError("detector extent exceeds the FP64 finite-ray range")


# Largest cell-layout copy (bytes) a plan keeps by default; larger fields
# gather from the linear array.
CELL_GATHER_MAX_BYTES = 32 << 30
# The copy is rebuilt at every call (about 36 bytes of traffic per cell) and
# saves under a picosecond per sample (full scan: 48 ms over 6.2e10 samples),
# so a plan takes it only when a call samples at least this many times per
# cell: the 360-view full scan does (~5800), one 512^2 view of a 512 x 512 x
# 139 chest grid does not (~7.5).
CELL_GATHER_MIN_SAMPLES_PER_CELL = 64


def gather_cells(decoded: Decoded) -> bool:
    """FP32 sampling gathers corners from a cell-layout copy of the field
    (one or two 16-byte loads per cell instead of four 8-byte loads) when a
    call samples at least CELL_GATHER_MIN_SAMPLES_PER_CELL times per cell and
    the copy fits CELL_GATHER_MAX_BYTES; DPT_AMD_GATHER=linear|cells
    overrides."""
    if decoded.spec.precision != "float32":
        return False
    override = os.environ.get("DPT_AMD_GATHER")
    if override == "linear":
        return False
    cells = _native.cell_count(decoded.grid.shape)
    element = 4 if decoded.spec.storage == "float32" else 2
    if 8 * element * cells > CELL_GATHER_MAX_BYTES:
        return False
    if override == "cells":
        return True
    views, rows, cols = decoded.image_shape
    return views * rows * cols * decoded.spec.samples_per_ray >= CELL_GATHER_MIN_SAMPLES_PER_CELL * cells


def prepare(value: str, device: int, stream: int, needs: int) -> Prepared:
    """Return the cached plan for this configuration, device, stream and flags."""
    key = (value, int(device), int(stream), int(needs))
    with _CACHE_LOCK:
        found = _CACHE.get(key)
        if found is not None:
            _CACHE.move_to_end(key)
            return found
        decoded = decode(value)
        _check_extent(decoded)
        if gather_cells(decoded):
            needs |= _native.GATHER_CELLS
        if decoded.spec.integration == "cell_gauss":
            needs |= _native.CELL_GAUSS
        plan = _native.Plan(
            decoded.grid,
            samples=decoded.spec.samples_per_ray,
            arithmetic=decoded.spec.precision,
            storage=decoded.spec.storage,
            needs=needs,
            device=int(device),
            **_native_geometry(decoded),
        )
        entry = Prepared(plan, decoded)
        _CACHE[key] = entry
        while len(_CACHE) > CACHE_LIMIT:
            _CACHE.popitem(last=False)
        return entry


def clear_cache() -> None:
    with _CACHE_LOCK:
        _CACHE.clear()


def consume(entry: Prepared, stream: int) -> None:
    """One readback of every view's status word; clear, then raise the first error."""
    codes = entry.plan.status(stream, clear=True)
    raise_view_statuses(codes, "projection")
