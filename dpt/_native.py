"""ctypes binding of the C ABI declared in include/dpt_amd.h.

The shared library is built in-tree by ``python -m dpt.build_native``.  The
product path has no fallback: if the library is missing, every projector call
raises ``NativeLibraryError``.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

from .contracts import ContractError

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libdpt_amd.so"

# every symbol include/dpt_amd.h declares
EXPORTS = (
    "dpt_plan_create",
    "dpt_plan_create_panels",
    "dpt_plan_destroy",
    "dpt_forward",
    "dpt_vjp",
    "dpt_ray_vjp",
    "dpt_pose_sensitivities",
    "dpt_reduce_rays",
    "dpt_adjoint_gather",
    "dpt_adjoint_gather_noalloc",
    "dpt_status",
    "dpt_convert_field",
    "dpt_cell_count",
    "dpt_convert_cells",
    "dpt_set_timing",
    "dpt_kernel_ms",
    "dpt_count_hit_rays",
    "dpt_plan_layout",
    "dpt_valu_probe",
    "dpt_compose_pose",
    "dpt_compose_pose_vjp",
    "dpt_last_error",
    "dpt_abi_version",
    "dpt_fit_create",
    "dpt_fit_destroy",
    "dpt_fit_step",
    "dpt_fit_last_error",
    "dpt_transport_create",
    "dpt_transport_destroy",
    "dpt_transport_trace",
    "dpt_transport_moments",
    "dpt_transport_status",
    "dpt_transport_last_error",
)

ARITH_FP32, ARITH_FP64 = 0, 1
STORE_F32, STORE_BF16, STORE_F16 = 0, 1, 2
# OR-ed into a storage code: the field is in the cell layout (dpt_convert_cells)
STORE_CELLS = 8
VOLUME, POSE, POSE_TANGENT, ACCUMULATE, NO_CHECKS, VOLUME_FP64 = 1, 2, 4, 8, 16, 32
# plan option (needs): FP32 sampling gathers from a cell-layout copy
GATHER_CELLS = 64
# plan option (needs, FP64 only): exact cell-Gauss integration
CELL_GAUSS = 128
# bytes of the fixed device workspace of dpt_adjoint_gather_noalloc
GATHER_NOALLOC_BYTES = 32 << 20

STORAGE_CODES = {"float32": STORE_F32, "bfloat16": STORE_BF16, "float16": STORE_F16}
# 16-bit field buffers carry DPT_HALF_PAD_BYTES (8) of readable padding
HALF_PAD_ELEMENTS = 4
ARITH_CODES = {"float32": ARITH_FP32, "float64": ARITH_FP64}


class NativeLibraryError(RuntimeError):
    """The HIP library is missing or a native call failed."""


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("object_to_grid", ctypes.c_double * 9),
        ("grid_origin", ctypes.c_double * 3),
    ]


class TransportModel(ctypes.Structure):
    """dpt_transport_model (include/dpt_amd.h)."""

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


class PointSource(ctypes.Structure):
    """dpt_point_source."""

    _fields_ = [
        ("source", ctypes.c_double * 3),
        ("corner", ctypes.c_double * 3),
        ("u", ctypes.c_double * 3),
        ("v", ctypes.c_double * 3),
        ("extent", ctypes.c_double * 2),
        ("weight", ctypes.c_double),
    ]


class HistoryRecords(ctypes.Structure):
    """dpt_history_records (device pointers, each optional)."""

    _fields_ = [
        ("pixel", ctypes.c_void_p),
        ("score", ctypes.c_void_p),
        ("energy", ctypes.c_void_p),
        ("events", ctypes.c_void_p),
        ("real_events", ctypes.c_void_p),
        ("status", ctypes.c_void_p),
    ]


_LIB = None
_LOCK = threading.Lock()


def lib():
    """Load the HIP library once; raise loudly when it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    with _LOCK:
        if _LIB is not None:
            return _LIB
        if not LIB_PATH.exists():
            raise NativeLibraryError(
                f"dpt HIP library not found at {LIB_PATH}; build it with "
                "`python -m dpt.build_native` (no CPU fallback exists)"
            )
        h = ctypes.CDLL(str(LIB_PATH))
        vp, i32, u32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64
        h.dpt_plan_create.argtypes = [
            ctypes.POINTER(GridDesc), i32, i32, i32, vp, vp, vp, i32, i32, i32, u32, i32,
            ctypes.POINTER(vp),
        ]
        h.dpt_plan_create.restype = ctypes.c_int
        h.dpt_plan_create_panels.argtypes = [
            ctypes.POINTER(GridDesc), i32, i32, i32, vp, vp, i32, i32, i32, u32, i32, ctypes.POINTER(vp),
        ]
        h.dpt_plan_create_panels.restype = ctypes.c_int
        h.dpt_plan_destroy.argtypes = [vp]
        h.dpt_plan_destroy.restype = ctypes.c_int
        h.dpt_forward.argtypes = [vp, vp, vp, vp, vp]
        h.dpt_forward.restype = ctypes.c_int
        h.dpt_vjp.argtypes = [vp, vp, vp, vp, vp, vp, u32, vp]
        h.dpt_vjp.restype = ctypes.c_int
        h.dpt_ray_vjp.argtypes = [vp, vp, vp, vp, vp, u32, vp]
        h.dpt_pose_sensitivities.argtypes = [vp, vp, vp, vp, vp, u32, vp]
        h.dpt_pose_sensitivities.restype = ctypes.c_int
        h.dpt_ray_vjp.restype = ctypes.c_int
        h.dpt_reduce_rays.argtypes = [vp, vp, vp, i32, vp, u32, vp]
        h.dpt_reduce_rays.restype = ctypes.c_int
        h.dpt_adjoint_gather.argtypes = [vp, vp, vp, vp, i32, u32, vp]
        h.dpt_adjoint_gather.restype = ctypes.c_int
        h.dpt_adjoint_gather_noalloc.argtypes = [vp, vp, vp, vp, i32, u32, vp]
        h.dpt_adjoint_gather_noalloc.restype = ctypes.c_int
        h.dpt_status.argtypes = [vp, vp, i32, vp]
        h.dpt_status.restype = ctypes.c_int
        h.dpt_convert_field.argtypes = [vp, vp, i64, i32, vp, vp]
        h.dpt_convert_field.restype = ctypes.c_int
        h.dpt_cell_count.argtypes = [i32, i32, i32]
        h.dpt_cell_count.restype = i64
        h.dpt_convert_cells.argtypes = [vp, vp, i32, i32, i32, i32, vp]
        h.dpt_convert_cells.restype = ctypes.c_int
        h.dpt_set_timing.argtypes = [vp, i32]
        h.dpt_set_timing.restype = ctypes.c_int
        h.dpt_kernel_ms.argtypes = [vp, i32]
        h.dpt_kernel_ms.restype = ctypes.c_double
        h.dpt_count_hit_rays.argtypes = [vp, vp, vp]
        h.dpt_plan_layout.argtypes = [vp, vp]
        h.dpt_plan_layout.restype = ctypes.c_int
        h.dpt_valu_probe.argtypes = [i32, vp, ctypes.POINTER(ctypes.c_double), vp]
        h.dpt_valu_probe.restype = ctypes.c_int
        h.dpt_compose_pose.argtypes = [vp, vp, i64, vp, vp, vp]
        h.dpt_compose_pose.restype = ctypes.c_int
        h.dpt_compose_pose_vjp.argtypes = [vp, vp, vp, i64, vp, vp, vp]
        h.dpt_compose_pose_vjp.restype = ctypes.c_int
        h.dpt_count_hit_rays.restype = i64
        h.dpt_last_error.argtypes = []
        h.dpt_last_error.restype = ctypes.c_char_p
        h.dpt_abi_version.argtypes = []
        h.dpt_abi_version.restype = i32
        u64, dbl = ctypes.c_uint64, ctypes.c_double
        h.dpt_transport_create.argtypes = [ctypes.POINTER(TransportModel), i32, ctypes.POINTER(vp)]
        h.dpt_transport_create.restype = ctypes.c_int
        h.dpt_transport_destroy.argtypes = [vp]
        h.dpt_transport_destroy.restype = ctypes.c_int
        h.dpt_transport_trace.argtypes = [
            vp, vp, ctypes.POINTER(PointSource), u64, u64, i64, dbl, ctypes.POINTER(HistoryRecords), vp, vp,
        ]
        h.dpt_transport_trace.restype = ctypes.c_int
        h.dpt_transport_moments.argtypes = [vp, vp, i64, vp, vp, vp]
        h.dpt_transport_moments.restype = ctypes.c_int
        h.dpt_transport_status.argtypes = [vp, vp, vp, i32, vp]
        h.dpt_transport_status.restype = ctypes.c_int
        h.dpt_transport_last_error.argtypes = []
        h.dpt_transport_last_error.restype = ctypes.c_char_p
        _LIB = h
        return h


def check(rc: int, what: str, error=None) -> None:
    if rc != 0:
        source = lib().dpt_last_error if error is None else error
        message = source().decode(errors="replace")
        if rc == -1:
            raise ContractError(f"{what}: {message}")
        raise NativeLibraryError(f"{what} failed ({rc}): {message}")


class Plan:
    """Owns one native dpt_plan (device metadata, status words, scratch)."""

    def __init__(
        self,
        grid,
        *,
        views: int,
        rows: int,
        cols: int,
        panels: np.ndarray | None,
        rays: tuple[np.ndarray, np.ndarray] | None,
        samples: int,
        arithmetic: str,
        storage: str,
        needs: int,
        device: int,
        panel_records: np.ndarray | None = None,
        panel_modes: np.ndarray | None = None,
    ) -> None:
        h = lib()
        desc = GridDesc()
        desc.nx, desc.ny, desc.nz = grid.xyz
        for i, value in enumerate(grid.object_to_grid_matrix()):
            desc.object_to_grid[i] = value
        for i in range(3):
            desc.grid_origin[i] = grid.origin_mm[i]
        self._keep = []
        panel_ptr = src_ptr = end_ptr = None
        if panel_records is not None:
            records = np.ascontiguousarray(panel_records, dtype=np.float64).reshape(views, 18)
            modes = np.ascontiguousarray(panel_modes, dtype=np.int32).reshape(views)
            self._keep += [records, modes]
            handle = ctypes.c_void_p()
            rc = h.dpt_plan_create_panels(
                ctypes.byref(desc), views, rows, cols, records.ctypes.data, modes.ctypes.data, samples,
                ARITH_CODES[arithmetic], STORAGE_CODES[storage], needs, device, ctypes.byref(handle),
            )
            check(rc, "dpt_plan_create_panels")
            self.handle = handle
            self.views, self.rows, self.cols = views, rows, cols
            self.arithmetic, self.storage = arithmetic, storage
            self.device = device
            self._codes = np.zeros(views, dtype=np.int32)
            return
        if panels is not None:
            arr = np.ascontiguousarray(panels, dtype=np.float64)
            self._keep.append(arr)
            panel_ptr = arr.ctypes.data
        else:
            src = np.ascontiguousarray(rays[0], dtype=np.float64)
            end = np.ascontiguousarray(rays[1], dtype=np.float64)
            self._keep += [src, end]
            src_ptr, end_ptr = src.ctypes.data, end.ctypes.data
        handle = ctypes.c_void_p()
        rc = h.dpt_plan_create(
            ctypes.byref(desc), views, rows, cols, panel_ptr, src_ptr, end_ptr, samples,
            ARITH_CODES[arithmetic], STORAGE_CODES[storage], needs, device, ctypes.byref(handle),
        )
        check(rc, "dpt_plan_create")
        self.handle = handle
        self.views, self.rows, self.cols = views, rows, cols
        self.arithmetic, self.storage = arithmetic, storage
        self.device = device
        self._codes = np.zeros(views, dtype=np.int32)

    def forward(self, field: int, poses: int, out: int, stream: int) -> None:
        check(lib().dpt_forward(self.handle, field, poses, out, stream), "dpt_forward")

    def vjp(self, field: int, poses: int, cotangent: int, grad_field: int | None,
            grad_pose: int | None, flags: int, stream: int) -> None:
        check(
            lib().dpt_vjp(self.handle, field, poses, cotangent, grad_field, grad_pose, flags, stream),
            "dpt_vjp",
        )

    def ray_vjp(self, field: int, poses: int, cotangent: int, out: int, flags: int, stream: int) -> None:
        check(lib().dpt_ray_vjp(self.handle, field, poses, cotangent, out, flags, stream), "dpt_ray_vjp")

    def pose_sensitivities(self, field: int, poses: int, cotangent: int, out: int, flags: int, stream: int) -> None:
        check(lib().dpt_pose_sensitivities(self.handle, field, poses, cotangent, out, flags, stream),
              "dpt_pose_sensitivities")

    def reduce_rays(self, rays: int, jacobians: int, components: int, out: int, flags: int, stream: int) -> None:
        check(lib().dpt_reduce_rays(self.handle, rays, jacobians, components, out, flags, stream), "dpt_reduce_rays")

    def adjoint_gather(self, poses: int, cotangent: int, out: int, out_precision: str, flags: int,
                       stream: int) -> None:
        check(
            lib().dpt_adjoint_gather(self.handle, poses, cotangent, out, ARITH_CODES[out_precision], flags, stream),
            "dpt_adjoint_gather",
        )

    def adjoint_gather_noalloc(self, poses: int, cotangent: int, out: int, out_precision: str, flags: int,
                               stream: int) -> None:
        """dpt_adjoint_gather on the library's fixed per-device workspace
        (allocates nothing; calls overlapping on one GPU race)."""
        check(
            lib().dpt_adjoint_gather_noalloc(self.handle, poses, cotangent, out, ARITH_CODES[out_precision], flags,
                                             stream),
            "dpt_adjoint_gather_noalloc",
        )

    def status(self, stream: int, clear: bool = True) -> np.ndarray:
        check(lib().dpt_status(self.handle, self._codes.ctypes.data, 1 if clear else 0, stream), "dpt_status")
        return self._codes.copy()

    def set_timing(self, enabled: bool) -> None:
        check(lib().dpt_set_timing(self.handle, 1 if enabled else 0), "dpt_set_timing")

    def kernel_ms(self, kind: int) -> float:
        return float(lib().dpt_kernel_ms(self.handle, kind))

    def count_hit_rays(self, poses: int, stream: int) -> int:
        value = int(lib().dpt_count_hit_rays(self.handle, poses, stream))
        if value < 0:
            check(value, "dpt_count_hit_rays")
        return value

    def layout(self) -> dict:
        """The launch layout the plan chose (brick entries, slab chunks,
        samples per chunk, views per VJP workgroup, cell-Gauss chunks per
        ray)."""
        out = np.zeros(5, dtype=np.int32)
        check(lib().dpt_plan_layout(self.handle, out.ctypes.data), "dpt_plan_layout")
        return {"brick_cap": int(out[0]), "slab_chunks": bool(out[1]), "chunk_samples": int(out[2]),
                "views_per_block": int(out[3]), "gauss_chunks": int(out[4])}

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().dpt_plan_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


POSE_JET = 84  # doubles per pose pair in the jet of dpt_compose_pose


def compose_pose(anchor: int, tangent: int, count: int, out: int, jet: int, stream: int) -> None:
    """anchor @ exp(hat(tangent)) for ``count`` pairs (device pointers)."""
    check(lib().dpt_compose_pose(anchor, tangent, count, out, jet, stream), "dpt_compose_pose")


def compose_pose_vjp(anchor: int, jet: int, grad: int, count: int, grad_anchor: int, grad_tangent: int,
                     stream: int) -> None:
    """Gradients of the chart from the jet of :func:`compose_pose`."""
    check(lib().dpt_compose_pose_vjp(anchor, jet, grad, count, grad_anchor, grad_tangent, stream),
          "dpt_compose_pose_vjp")


def valu_probe(kind: int, scratch: int, stream: int) -> float:
    """Wave64 VALU instructions per second of 16 independent v_fma_f32
    (kind 0) or v_add_u32 (kind 1) chains per lane at 8 waves per SIMD;
    scratch: device buffer of 2048 * 256 floats."""
    rate = ctypes.c_double()
    check(lib().dpt_valu_probe(kind, scratch, ctypes.byref(rate), stream), "dpt_valu_probe")
    return float(rate.value)


def cell_count(shape) -> int:
    """Cells of a (nz, ny, nx) grid in the cell layout: prod(max(n - 1, 1))."""
    nz, ny, nx = (int(n) for n in shape)
    count = lib().dpt_cell_count(nx, ny, nz)
    if count < 0:
        check(int(count), "dpt_cell_count")
    return int(count)


def convert_cells(src: int, dst: int, shape, storage: str, stream: int) -> None:
    """Rewrite a linear (nz, ny, nx) field into the cell layout (same dtype)."""
    nz, ny, nx = (int(n) for n in shape)
    check(lib().dpt_convert_cells(src, dst, nx, ny, nz, STORAGE_CODES[storage], stream), "dpt_convert_cells")


def convert_field(src: int, dst: int, count: int, storage: str, status: int | None, stream: int) -> None:
    check(
        lib().dpt_convert_field(src, dst, count, STORAGE_CODES[storage], status, stream),
        "dpt_convert_field",
    )
