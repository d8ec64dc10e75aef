// Deterministic adjoint gather: launch interface shared by gather.hip (the
// kernels) and projector.hip (the plan-owning C entry point).
#pragma once

#include "device_common.h"

namespace dpt {

// One ray of the gather, prepared once per call: the FP64 ray of the forward
// march (kernels/projection.py:79-121) with its clipped interval, sample
// step and volume factor (kernels/projection_adjoint.py:94-157), plus the
// reciprocals and the box widening the culling uses.  vf == 0 marks a ray
// that adds nothing (invalid, zero cotangent or empty interval).
struct alignas(16) GatherRay {
    double o[3], d[3], inv[3];  // grid-frame origin, direction, 1/direction (0 where direction is 0)
    double lower, step, inv_step, vf;
    double margin;              // support-box widening for this ray's sample points, grid units
};
static_assert(sizeof(GatherRay) == 112, "a ray record is seven 16-byte loads");

// Per-view projection of grid points onto the panel's (column, row)
// coordinates for the detector culling: a point x lies on the ray of pixel
// (c, r) when (x - s).B = c (x - s).A and (x - s).C = r (x - s).A, with
// A = U x V, B = V x D0, C = D0 x U in grid coordinates (U, V: column and
// row steps, D0: direction of pixel (0, 0), s: source), signs chosen so that
// (x - s).A > 0 in front of the source.
struct alignas(16) GatherView {
    double s[3], a[3], b[3], c[3];
    double norm_a;  // |A|
    double ratio;   // (|B| + |C|) / |A|, the scale of the pixel-coordinate rounding
    double margin;  // support-box widening for this view's rays, grid units
    int ok;          // 0: cull nothing (explicit rays, cylindrical or degenerate panels)
    int affine;      // 1: parallel panel, pixel coordinates (x - s).b and (x - s).c with no denominator
    int cols, rows;  // detector columns and rows of the view; ray (row, col) is at row * cols + col
};

// Enqueue the gather: guard reset, cotangent/pose checks (checks != 0), view
// and ray preparation, then one thread per voxel.  out is double (out_fp64)
// or float; the guard word suppresses every write when a check failed.
// panel_meta / panel_modes: the device panel records of a plan made from
// panel records (views x 18, views) with image_cols detector columns, or
// null.
hipError_t launch_gather(const Params &P, const double *cotangent, GatherRay *rays, GatherView *views, int *guard,
                         void *out, bool out_fp64, bool accumulate, bool checks, const double *panel_meta,
                         const int *panel_modes, int image_cols, hipStream_t stream);

}  // namespace dpt
