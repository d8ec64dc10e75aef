// Deterministic adjoint gather for gfx950: out[v] = sum_p A[p, v] c[p] in
// FP64 with one thread per voxel, so every voxel's sum has a fixed order and
// the result is the same bit for bit on every run (the ray-driven volume
// adjoint accumulates with atomics and is not).
//
// Restates kernels/projection_adjoint.py:73-170 (get_gather, midpoint):
// for each voxel, views in order, pixels row-major, samples in order and the
// 8 cell corners in order, add scaled_corner_weight(fraction, corner,
// volume_factor) (kernels/projection.py:396-404, ((vf * w_x) * w_y) * w_z,
// no FP32 rounding) for every corner whose clamped index is the voxel.
// Compiled with -ffp-contract=off and written in the reference's operation
// order, so the sums equal the CPU restatement's bit for bit.
//
// Culling (the role of kernels/projection_adjoint.py:330-597, designed
// here independently): a voxel can only be a corner of a sample whose
// clamped cell coordinates lie in [i-1, i+1] per axis (to the support
// boundary -0.5 / n-0.5 at the grid edges).  The kernel visits only the
// pixels whose rays can reach that box, found by projecting the box's 8
// corners onto the panel from the source, and only the samples of each ray
// whose points can fall in it, found by slab clipping.  Both enclosures are
// widened by margins many orders of magnitude above the FP64 rounding of
// the sample points, so every skipped sample adds exactly nothing (no corner
// of its cell is the voxel) and the sums are those of the full loop.  A view
// whose panel is nearly edge-on to part of the box falls back to all pixels.
// Panel records: flat cone panels project as above and parallel panels
// affinely along the beam; explicit rays and cylindrical panels visit every
// pixel.
#include "gather.h"

#include <algorithm>
#include <cmath>

namespace dpt {

namespace {

// Relative support-box widening (grid units per unit of coordinate
// magnitude): the rounding of a sample point is below 1e-15 of the
// magnitudes involved.
constexpr double kBoxMargin = 1e-9;
// Relative widening of a projected pixel range and of a ray parameter range.
constexpr double kPixelMargin = 1e-9;
constexpr double kParamMargin = 1e-12;
// Minimum (x - s).A / (|x - s| |A|) at every box corner for the projection
// to be used (about 0.06 degrees off the plane through the source parallel
// to the panel); otherwise the view visits every pixel.
constexpr double kFacing = 1e-3;
// Steps below this (grid units) visit every sample of the ray.
constexpr double kMinStep = 1e-12;

__device__ __forceinline__ void grid_vector(const Params &P, const double *pose, const double *x, double *out) {
    double od[3];
    for (int i = 0; i < 3; ++i) od[i] = pose[i] * x[0] + pose[3 + i] * x[1] + pose[6 + i] * x[2];
    for (int i = 0; i < 3; ++i) out[i] = P.g[3 * i] * od[0] + P.g[3 * i + 1] * od[1] + P.g[3 * i + 2] * od[2];
}

// Grid coordinates of a world point under the view's pose (the ray-origin
// arithmetic of kernels/projection.py:79-102)
__device__ __forceinline__ void grid_point(const Params &P, const double *pose, const double *x, double *out) {
    const double w[3] = {x[0] - pose[9], x[1] - pose[10], x[2] - pose[11]};
    double rel[3];
    for (int i = 0; i < 3; ++i) rel[i] = pose[i] * w[0] + pose[3 + i] * w[1] + pose[6 + i] * w[2] - P.go[i];
    for (int i = 0; i < 3; ++i) out[i] = P.g[3 * i] * rel[0] + P.g[3 * i + 1] * rel[1] + P.g[3 * i + 2] * rel[2];
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *out) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ double max_abs3(const double *a) {
    return fmax(fabs(a[0]), fmax(fabs(a[1]), fabs(a[2])));
}

__global__ void clear_guard_kernel(int *guard) { *guard = 0; }

// Cotangent finite (NumericalError) and poses rigid (ContractError), as the
// reference validates before launching (python/dpt/projection.py:1337-1363);
// a failure also raises the guard so that the gather writes nothing.
__global__ __launch_bounds__(kBlock) void gather_checks_kernel(const double *__restrict__ poses, int views,
                                                               const double *__restrict__ cotangent, int64_t pixels,
                                                               int *status, int *guard) {
    const int64_t first = (int64_t)blockIdx.x * kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
    bool bad = false;
    for (int64_t i = first; i < pixels; i += stride) bad |= !isfinite(cotangent[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) {
        atomicOr(&status[0], DPT_STATUS_RANGE);
        atomicOr(guard, 1);
    }
    for (int64_t v = first; v < views; v += stride)
        if (!pose_is_rigid(poses + 12 * v)) {
            atomicOr(&status[v], DPT_STATUS_POSE);
            atomicOr(guard, 1);
        }
}

// Parallel panel (segments generated from its record, geometry.py:349-354
// of python/dpt): a grid point x lies on the segment line of pixel (c, r)
// when x - E0 = c U + r V + t D (E0 = pixel (0,0) endpoint, U, V = column
// and row steps, D = beam, all in grid coordinates), so
// c = (x - E0).(V x D) / det and r = (x - E0).(D x U) / det with
// det = U.(V x D): affine, its extremes over a box are at the corners.  A
// beam within about 0.06 degrees of the panel plane (|det| below
// kFacing |U| |V| |D|) visits every pixel.
__device__ __forceinline__ void parallel_view(const Params &P, const double *pose, const double *m, int cols, int rows,
                                              GatherView &gv) {
    double e0[3], u[3], vv[3], d[3], uw[3], vw[3], dw[3];
    for (int k = 0; k < 3; ++k) {
        uw[k] = m[12] * m[6 + k];
        vw[k] = m[13] * m[9 + k];
        dw[k] = m[14] * m[k];
    }
    grid_point(P, pose, m + 3, e0);
    grid_vector(P, pose, uw, u);
    grid_vector(P, pose, vw, vv);
    grid_vector(P, pose, dw, d);
    double b[3], c[3], vd[3];
    cross3(vv, d, vd);
    cross3(d, u, c);
    const double det = dot3(u, vd);
    const double volume = sqrt(dot3(u, u)) * sqrt(dot3(vv, vv)) * sqrt(dot3(d, d));
    if (!(isfinite(det) && isfinite(volume) && fabs(det) > kFacing * volume)) return;
    for (int k = 0; k < 3; ++k) {
        gv.s[k] = e0[k];
        gv.b[k] = vd[k] / det;
        gv.c[k] = c[k] / det;
        b[k] = gv.b[k];
    }
    // the segment points' rounding scales with the world coordinates before
    // the pose and grid maps as well as with the grid coordinates
    double gmax = 0.0;
    for (int k = 0; k < 9; ++k) gmax = fmax(gmax, fabs(P.g[k]));
    const double world = max_abs3(m + 3) + max_abs3(pose + 9) + max_abs3(P.go) + m[14] * max_abs3(m) +
                         (double)(cols - 1) * m[12] * max_abs3(m + 6) + (double)(rows - 1) * m[13] * max_abs3(m + 9);
    const double reach = max_abs3(d) + (double)(cols - 1) * max_abs3(u) + (double)(rows - 1) * max_abs3(vv);
    const double extent = (double)max(P.nx, max(P.ny, P.nz));
    gv.margin = kBoxMargin * (1.0 + max_abs3(e0) + reach + 3.0 * gmax * world);
    // |x - E0| <= |x| + |E0| over the widened boxes: bounds the rounding of
    // the pixel coordinates together with that of b and c
    gv.ratio = (sqrt(dot3(b, b)) + sqrt(dot3(gv.c, gv.c))) * (1.0 + extent + max_abs3(e0) + reach);
    gv.norm_a = 1.0;
    gv.affine = 1;
    gv.ok = isfinite(gv.ratio) && isfinite(gv.margin);
}

// Flat cone panel pn (source, pixel (0,0) centre, column step, row step;
// world coordinates): the projective pixel coordinates through the source.
__device__ __forceinline__ void cone_view(const Params &P, const double *pose, const double *pn, int cols, int rows,
                                          GatherView &gv) {
    World w;
    for (int i = 0; i < 3; ++i) {
        w.wsr[i] = pn[i] - pose[9 + i];
        w.wd[i] = pn[3 + i] - pn[i];
    }
    for (int i = 0; i < 3; ++i) {
        w.os[i] = pose[i] * w.wsr[0] + pose[3 + i] * w.wsr[1] + pose[6 + i] * w.wsr[2];
        w.od[i] = pose[i] * w.wd[0] + pose[3 + i] * w.wd[1] + pose[6 + i] * w.wd[2];
    }
    Ray r;
    ray_grid(P, w, r);
    double u[3], vv[3], a[3], b[3], c[3];
    grid_vector(P, pose, pn + 6, u);
    grid_vector(P, pose, pn + 9, vv);
    cross3(u, vv, a);
    cross3(vv, r.d, b);
    cross3(r.d, u, c);
    const double det = dot3(r.d, a);
    if (r.valid && isfinite(det) && det != 0.0) {
        const double sign = det > 0.0 ? 1.0 : -1.0;
        for (int i = 0; i < 3; ++i) {
            gv.s[i] = r.o[i];
            gv.a[i] = sign * a[i];
            gv.b[i] = sign * b[i];
            gv.c[i] = sign * c[i];
        }
        gv.norm_a = sqrt(dot3(a, a));
        gv.ratio = (sqrt(dot3(b, b)) + sqrt(dot3(c, c))) / gv.norm_a;
        const double reach = max_abs3(r.d) + (double)(cols - 1) * max_abs3(u) + (double)(rows - 1) * max_abs3(vv);
        gv.margin = kBoxMargin * (1.0 + max_abs3(r.o) + reach);
        gv.ok = isfinite(gv.ratio) && isfinite(gv.margin) && gv.norm_a > 0.0 && isfinite(dot3(gv.b, gv.b)) &&
                isfinite(dot3(gv.c, gv.c));
    }
}

// One thread per view: the panel projection used by the detector culling.
__global__ void gather_views_kernel(Params P, const double *__restrict__ meta, const int *__restrict__ modes,
                                    int image_rows, int image_cols, GatherView *__restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.views) return;
    GatherView gv;
    for (int i = 0; i < 3; ++i) gv.s[i] = gv.a[i] = gv.b[i] = gv.c[i] = 0.0;
    gv.norm_a = gv.ratio = gv.margin = 0.0;
    gv.ok = 0;
    gv.affine = 0;
    gv.cols = P.bundle && meta != nullptr ? image_cols : P.cols;
    gv.rows = P.bundle && meta != nullptr ? image_rows : P.rows;
    if (P.bundle && meta != nullptr && modes[v] == 1) {
        parallel_view(P, P.poses + 12 * (size_t)v, meta + 18 * (size_t)v, image_cols, image_rows, gv);
    } else if (P.bundle && meta != nullptr && modes[v] == 0) {
        // flat cone panel from its record (source, pixel (0,0), u, v, pitches)
        const double *m = meta + 18 * (size_t)v;
        double pn[12];
        for (int k = 0; k < 3; ++k) {
            pn[k] = m[k];
            pn[3 + k] = m[3 + k];
            pn[6 + k] = m[12] * m[6 + k];
            pn[9 + k] = m[13] * m[9 + k];
        }
        cone_view(P, P.poses + 12 * (size_t)v, pn, image_cols, image_rows, gv);
    } else if (!P.bundle) {
        cone_view(P, P.poses + 12 * (size_t)v, P.panels + 12 * (size_t)v, P.cols, P.rows, gv);
    }
    out[v] = gv;
}

// One thread per ray: setup, validity, interval and volume factor
// (kernels/projection_adjoint.py:89-157); rays flagged and skipped as there.
__global__ __launch_bounds__(kBlock) void gather_rays_kernel(Params P, const double *__restrict__ cotangent,
                                                             GatherRay *__restrict__ rays) {
    const int64_t pixels = (int64_t)P.rows * P.cols, total = (int64_t)P.views * pixels;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock) {
        const int view = (int)(i / pixels);
        const int pixel = (int)(i - (int64_t)view * pixels);
        const int row = pixel / P.cols, col = pixel - row * P.cols;
        const double *pose = P.poses + 12 * (size_t)view;
        World w;
        ray_world(P, pose, view, row, col, w);
        Ray r;
        ray_grid(P, w, r);
        GatherRay g;
        for (int a = 0; a < 3; ++a) g.o[a] = g.d[a] = g.inv[a] = 0.0;
        g.lower = g.step = g.inv_step = g.vf = g.margin = 0.0;
        if (!r.valid) {
            flag(P, view, DPT_STATUS_RANGE);
        } else {
            const double seed = cotangent[i];
            const Interval64 iv = interval64(P, r.o, r.d);
            const double span = iv.upper - iv.lower;
            if (seed != 0.0 && span > 0.0) {
                const double factor = seed * r.len;
                const double step = span / (double)P.samples;
                double vf = factor * span / (double)P.samples;
                if (!isfinite(factor)) vf = product4(seed, r.len, span, 1.0 / (double)P.samples);
                // |pos| <= 1, so points are finite unless the ray's
                // coordinates approach the FP64 range: only then test them.
                if (fmax(max_abs3(r.o), max_abs3(r.d)) > 1e300) {
                    bool bad = false;
                    for (int k = 0; k < P.samples; ++k) {
                        const double pos = iv.lower + ((double)k + 0.5) * step;
                        const double p[3] = {r.o[0] + pos * r.d[0], r.o[1] + pos * r.d[1], r.o[2] + pos * r.d[2]};
                        bad |= !finite3d(p);
                    }
                    if (bad) flag(P, view, DPT_STATUS_RANGE);
                }
                for (int a = 0; a < 3; ++a) {
                    g.o[a] = r.o[a];
                    g.d[a] = r.d[a];
                    g.inv[a] = 1.0 / r.d[a];  // +-inf for a zero component
                }
                g.lower = iv.lower;
                g.step = step;
                g.inv_step = 1.0 / step;
                g.vf = vf;
                // the rounding of a sample point is below 1e-15 of |o| + |d|
                double reach = 0.0;
                for (int a = 0; a < 3; ++a) reach = fmax(reach, fabs(r.o[a]) + fabs(r.d[a]));
                g.margin = kBoxMargin * (1.0 + reach);
            }
        }
        rays[i] = g;
    }
}

struct Voxel {
    int i[3];
    double c[3];          // i as doubles
    double lo[3], hi[3];  // clamped-cell support: lower cell coordinate in [i-1, i] per axis
};

// Pixel window [c0, c1] x [r0, r1] of one view whose rays can reach the
// voxel's support box; false when the projection cannot be trusted (the
// caller then visits every pixel).
__device__ __forceinline__ bool pixel_window(const Params &P, const GatherView &gv, const Voxel &vx, int &c0,
                                             int &c1, int &r0, int &r1) {
    // Corner q of the widened box is w_lo + (q&1) e_x + (q>>1&1) e_y +
    // (q>>2&1) e_z relative to the source; its three dot products are sums
    // of per-axis terms (rounding differences are far inside the margins).
    if (gv.affine) {
        // parallel panel: affine pixel coordinates, no facing condition
        double cmin = INFINITY, cmax = -INFINITY, rmin = INFINITY, rmax = -INFINITY;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double w[3];
            for (int a = 0; a < 3; ++a)
                w[a] = (((q >> a) & 1) ? vx.hi[a] + gv.margin : vx.lo[a] - gv.margin) - gv.s[a];
            const double c = dot3(w, gv.b), r = dot3(w, gv.c);
            cmin = fmin(cmin, c);
            cmax = fmax(cmax, c);
            rmin = fmin(rmin, r);
            rmax = fmax(rmax, r);
        }
        if (!isfinite(cmin) || !isfinite(cmax) || !isfinite(rmin) || !isfinite(rmax)) return false;
        const double mc = kPixelMargin * (1.0 + fmax(fabs(cmin), fabs(cmax)) + gv.ratio);
        const double mr = kPixelMargin * (1.0 + fmax(fabs(rmin), fabs(rmax)) + gv.ratio);
        c0 = max((int)ceil(fmin(fmax(cmin - mc, -1.0), (double)gv.cols)), 0);
        c1 = min((int)floor(fmax(fmin(cmax + mc, (double)gv.cols), -1.0)), gv.cols - 1);
        r0 = max((int)ceil(fmin(fmax(rmin - mr, -1.0), (double)gv.rows)), 0);
        r1 = min((int)floor(fmax(fmin(rmax + mr, (double)gv.rows), -1.0)), gv.rows - 1);
        return true;
    }
    double den0 = 0.0, cn0 = 0.0, rn0 = 0.0, reach = 0.0;
    double da[3], db[3], dc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double lo = vx.lo[a] - gv.margin - gv.s[a];
        const double e = (vx.hi[a] - vx.lo[a]) + 2.0 * gv.margin;
        den0 += lo * gv.a[a];
        cn0 += lo * gv.b[a];
        rn0 += lo * gv.c[a];
        da[a] = e * gv.a[a];
        db[a] = e * gv.b[a];
        dc[a] = e * gv.c[a];
        reach += fmax(fabs(lo), fabs(lo + e));  // |w|_1 bound over the box (>= |w|_2)
    }
    const double threshold = kFacing * gv.norm_a * reach;
    double cmin = INFINITY, cmax = -INFINITY, rmin = INFINITY, rmax = -INFINITY;
    bool good = true;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const double den = den0 + ((q & 1) ? da[0] : 0.0) + ((q & 2) ? da[1] : 0.0) + ((q & 4) ? da[2] : 0.0);
        const double cn = cn0 + ((q & 1) ? db[0] : 0.0) + ((q & 2) ? db[1] : 0.0) + ((q & 4) ? db[2] : 0.0);
        const double rn = rn0 + ((q & 1) ? dc[0] : 0.0) + ((q & 2) ? dc[1] : 0.0) + ((q & 4) ? dc[2] : 0.0);
        good = good && den > threshold;
        // hardware reciprocal refined by two Newton steps (relative error far
        // below the pixel margin; only used when den is positive and normal)
        double inv = __builtin_amdgcn_rcp(den);
        inv = fma(inv, fma(-den, inv, 1.0), inv);
        inv = fma(inv, fma(-den, inv, 1.0), inv);
        const double c = cn * inv, r = rn * inv;
        cmin = fmin(cmin, c);
        cmax = fmax(cmax, c);
        rmin = fmin(rmin, r);
        rmax = fmax(rmax, r);
    }
    if (!good || !isfinite(cmin) || !isfinite(cmax) || !isfinite(rmin) || !isfinite(rmax)) return false;
    const double mc = kPixelMargin * (1.0 + fmax(fabs(cmin), fabs(cmax)) + gv.ratio);
    const double mr = kPixelMargin * (1.0 + fmax(fabs(rmin), fabs(rmax)) + gv.ratio);
    // clamped to [-1, n] before the integer conversion
    c0 = (int)ceil(fmin(fmax(cmin - mc, -1.0), (double)gv.cols));
    c1 = (int)floor(fmax(fmin(cmax + mc, (double)gv.cols), -1.0));
    r0 = (int)ceil(fmin(fmax(rmin - mr, -1.0), (double)gv.rows));
    r1 = (int)floor(fmax(fmin(rmax + mr, (double)gv.rows), -1.0));
    c0 = max(c0, 0);
    r0 = max(r0, 0);
    c1 = min(c1, gv.cols - 1);
    r1 = min(r1, gv.rows - 1);
    return true;
}

// Sample indices [k0, k1] of ray g whose points can lie in the voxel's
// support box; false when none can.  Branch-free slab clip of the widened
// box (lanes of a wave test different rays, so every lane runs the same
// instructions).  A zero direction component has the reciprocal +-inf, so
// its slab is by itself unbounded (the constant coordinate inside: t from
// -inf to +inf) or empty (outside: both t at the same infinity); a widened
// bound exactly at the coordinate gives 0 * inf = NaN, which fmin / fmax
// ignore, leaving the range empty -- right, as that coordinate lies the
// margin outside the unwidened box.
__device__ __forceinline__ bool sample_window(const Params &P, const GatherRay &g, const Voxel &vx, int &k0,
                                              int &k1) {
    double tlo = -INFINITY, thi = INFINITY;
    const double m = g.margin;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double t1 = ((vx.lo[a] - m) - g.o[a]) * g.inv[a], t2 = ((vx.hi[a] + m) - g.o[a]) * g.inv[a];
        tlo = fmax(tlo, fmin(t1, t2));
        thi = fmin(thi, fmax(t1, t2));
    }
    tlo -= kParamMargin * (1.0 + fabs(tlo));
    thi += kParamMargin * (1.0 + fabs(thi));
    // an empty range (infinite ends give NaN here and also count as empty)
    const bool meets = tlo <= thi;
    // sample k lies in [tlo, thi] when kf0 <= k <= kf1; the box and t
    // margins already exceed the point rounding by orders of magnitude, the
    // slack only absorbs the rounding of kf itself
    const double slack = 1e-6 + 1e-14 * g.inv_step;
    double kf0 = (tlo - g.lower) * g.inv_step - 0.5 - slack, kf1 = (thi - g.lower) * g.inv_step - 0.5 + slack;
    const bool exact = g.step > kMinStep && !isnan(kf0) && !isnan(kf1);
    const double top = (double)P.samples + 1.0;
    kf0 = fmin(fmax(kf0, -2.0), top);
    kf1 = fmin(fmax(kf1, -2.0), top);
    const int a0 = max(0, (int)ceil(kf0)), a1 = min(P.samples - 1, (int)floor(kf1));
    k0 = exact ? a0 : 0;
    k1 = exact ? a1 : P.samples - 1;
    return meets && (!exact || a0 <= a1);
}

// One ray record in seven 16-byte loads, issued together.
__device__ __forceinline__ GatherRay load_ray(const GatherRay *r) {
    const double2 *q = reinterpret_cast<const double2 *>(r);
    const double2 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5], q6 = q[6];
    GatherRay g;
    g.o[0] = q0.x;
    g.o[1] = q0.y;
    g.o[2] = q1.x;
    g.d[0] = q1.y;
    g.d[1] = q2.x;
    g.d[2] = q2.y;
    g.inv[0] = q3.x;
    g.inv[1] = q3.y;
    g.inv[2] = q4.x;
    g.lower = q4.y;
    g.step = q5.x;
    g.inv_step = q5.y;
    g.vf = q6.x;
    g.margin = q6.y;
    return g;
}

// How a voxel's corner test is evaluated (all three give the reference's
// addends bit for bit):
//   kGeneral  any grid: the 8 corners of the sample's clamped cell in order,
//             each compared with the voxel (clamped duplicates on singleton
//             axes are added once per corner, as in the reference);
//   kSimple   every axis has >= 2 samples: at most one corner is the voxel;
//   kMasked   every axis has >= 2 samples and the ray's factor is finite:
//             the product of three per-axis weights is added for every
//             sample, a weight being 0 where the reference finds no match.
//             On an axis of MASK (the voxel coordinate i in [2, n-3] for
//             every lane of the wave) a matching cell lies in [1, n-2),
//             where the support test and the clamps are the identity, and a
//             point outside the support or clamped gives a lower corner of
//             -1 or less, 0 or n-2, none of which can match.  With
//             e = p - i the lower corner floor(p) is i - 1 or i exactly when
//             -1 <= e < 1, and the reference's weight (1 - f for lower
//             corner i, f = p - (i - 1) for i - 1) is 1 - |e| bit for bit:
//             p - i and p - (i - 1) are exact there (Sterbenz, i >= 2), so
//             1 - |e| is the representable f or the same 1 - f; every other
//             point has max(1 - |e|, 0) = 0 (|e| >= 1, or a point just
//             below i - 1 whose e rounds to -1).  Other axes take the
//             support test, clamps and match of kSimple.  A sample without
//             a match adds a signed zero, which leaves the sum unchanged (a
//             running sum that starts at +0 is never -0).
constexpr int kGeneral = 0, kSimple = 1, kMasked = 2;

// Add the contributions of ray g's samples k0..k1 to the voxel's total, in
// sample then corner order (kernels/projection_adjoint.py:158-163 with
// point_contribution, :33-44).
template <int MODE, int MASK = 0>
__device__ __forceinline__ double add_ray(const Params &P, const GatherRay &g, const Voxel &vx, int k0, int k1,
                                          double total) {
    const int n[3] = {P.nx, P.ny, P.nz};
    const int cap[3] = {P.capx, P.capy, P.capz};
    const double vf = g.vf;
    if constexpr (MODE == kMasked) {
        double kc = (double)k0 + 0.5;  // k + 0.5, exact
        for (int k = k0; k <= k1; ++k, kc += 1.0) {
            const double pos = g.lower + kc * g.step;
            double w[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double p = g.o[a] + pos * g.d[a];
                if ((MASK >> a) & 1) {
                    w[a] = fmax(1.0 - fabs(p - vx.c[a]), 0.0);
                } else {
                    // kernels/projection.py:124-134, 199-214
                    const bool inside = p >= -0.5 && p <= (double)n[a] - 0.5;
                    const double c = fmin(fmax(p, 0.0), (double)(n[a] - 1));
                    const int l = min((int)floor(c), cap[a]);
                    const double f = c - (double)l;
                    w[a] = !inside ? 0.0 : l == vx.i[a] ? 1.0 - f : l + 1 == vx.i[a] ? f : 0.0;
                }
            }
            total += ((vf * w[0]) * w[1]) * w[2];
        }
        return total;
    }
    for (int k = k0; k <= k1; ++k) {
        const double pos = g.lower + ((double)k + 0.5) * g.step;
        double p[3];
        for (int a = 0; a < 3; ++a) p[a] = g.o[a] + pos * g.d[a];
        // kernels/projection.py:124-134 (inside_support; NaN fails every test)
        if (!(p[0] >= -0.5 && p[1] >= -0.5 && p[2] >= -0.5 && p[0] <= (double)n[0] - 0.5 &&
              p[1] <= (double)n[1] - 0.5 && p[2] <= (double)n[2] - 0.5))
            continue;
        int l[3];
        double f[3];
        // kernels/projection.py:199-214 (cell_lower, cell_fraction)
        for (int a = 0; a < 3; ++a) {
            const double c = fmin(fmax(p[a], 0.0), (double)(n[a] - 1));
            l[a] = min((int)floor(c), cap[a]);
            f[a] = c - (double)l[a];
        }
        if constexpr (MODE == kSimple) {
            bool hit = true;
            double w[3];
            for (int a = 0; a < 3; ++a) {
                hit = hit && (l[a] == vx.i[a] || l[a] + 1 == vx.i[a]);
                w[a] = l[a] == vx.i[a] ? 1.0 - f[a] : f[a];
            }
            if (hit) total += ((vf * w[0]) * w[1]) * w[2];
        } else {
            bool m0[3], m1[3];
            for (int a = 0; a < 3; ++a) {
                m0[a] = l[a] == vx.i[a];
                m1[a] = min(l[a] + 1, n[a] - 1) == vx.i[a];
            }
#pragma unroll
            for (int corner = 0; corner < 8; ++corner) {
                const int bx = corner & 1, by = (corner >> 1) & 1, bz = (corner >> 2) & 1;
                if ((bx ? m1[0] : m0[0]) && (by ? m1[1] : m0[1]) && (bz ? m1[2] : m0[2]))
                    total += ((vf * (bx ? f[0] : 1.0 - f[0])) * (by ? f[1] : 1.0 - f[1])) * (bz ? f[2] : 1.0 - f[2]);
            }
        }
    }
    return total;
}

// The voxel's FP64 sum over every view, pixel window and sample window.
template <int MODE, int MASK = 0>
__device__ __forceinline__ double gather_voxel(const Params &P, const GatherRay *__restrict__ rays,
                                               const GatherView *__restrict__ views, const Voxel &vx) {
    const int64_t pixels = (int64_t)P.rows * P.cols;
    double total = 0.0;
    for (int view = 0; view < P.views; ++view) {
        const GatherView gv = views[view];
        int c0 = 0, c1 = gv.cols - 1, r0 = 0, r1 = gv.rows - 1;
        if (gv.ok && !pixel_window(P, gv, vx, c0, c1, r0, r1)) {
            c0 = 0;
            c1 = gv.cols - 1;
            r0 = 0;
            r1 = gv.rows - 1;
        }
        // the window's pixels row-major in one loop, each ray record loaded
        // one pixel ahead so that its latency overlaps the current ray
        const int width = c1 - c0 + 1;
        const int count = (c1 >= c0 && r1 >= r0) ? width * (r1 - r0 + 1) : 0;
        const GatherRay *line = rays + (int64_t)view * pixels + (int64_t)r0 * gv.cols;
        int col = c0 - 1;
        for (int j = 0; j < count; ++j) {
            if (++col > c1) {
                col = c0;
                line += gv.cols;
            }
            const GatherRay g = load_ray(line + col);
            int k0, k1;
            const bool reach = sample_window(P, g, vx, k0, k1);
            if (g.vf != 0.0 && reach) {
                if (MODE == kMasked && !isfinite(g.vf))
                    total = add_ray<kSimple>(P, g, vx, k0, k1, total);
                else
                    total = add_ray<MODE, MASK>(P, g, vx, k0, k1, total);
            }
        }
    }
    return total;
}

// Brick order of the launch.  Voxels within two samples of a grid face get
// every sample of the short rays that graze that face or its edges -- tens
// of times the samples of an interior voxel -- so the bricks on the surface
// of the brick grid go first, one to each XCD in turn, and run beside the
// rest instead of after it.  The interior bricks follow, each XCD taking a
// contiguous range so that its L2 keeps the ray records of neighbouring
// bricks (workgroup b runs on XCD b % 8; surface is padded to a multiple of
// 8 so that interior workgroups keep that correspondence).
//
// The voxels along the grid's 12 edges (two coordinates within two samples
// of a face) collect the most: every sample of every ray that clips that
// edge, up to a hundred times an interior voxel.  They go before everything
// else, each wave taking 64 consecutive voxels along one edge line: those
// voxels' pixel windows advance through the same columns (or rows) in step,
// so their long windows coincide in the wave instead of queueing one after
// another.  The lanes of surface bricks that fall on an edge line are idle.
struct BrickLayout {
    int edges;             // edge waves (0: no edge section, a grid axis below 4 samples)
    int zw, xw, yw;        // waves per edge line along z, x, y
    int edge_groups;       // workgroups of the edge section, a multiple of kXcds
    int nx, ny, nz;        // bricks per axis
    int surface;           // bricks on the surface of the brick grid
    int surface_padded;    // rounded up to a multiple of kXcds
    int interior;          // (nx - 2)(ny - 2)(nz - 2), or 0
    int chunk;             // interior bricks per XCD
};

BrickLayout brick_layout(const Params &P) {
    BrickLayout L;
    L.edges = L.zw = L.xw = L.yw = L.edge_groups = 0;
    if (P.nx >= 4 && P.ny >= 4 && P.nz >= 4) {
        // 16 lines per edge family (4 x 4 band coordinates); lines along x
        // and y leave out their band ends, which lie on the lines along z
        // or at a corner
        L.zw = (P.nz + 63) / 64;
        L.xw = (P.nx - 4 + 63) / 64;
        L.yw = (P.ny - 4 + 63) / 64;
        L.edges = 16 * (L.zw + L.xw + L.yw);
        L.edge_groups = ((L.edges + 3) / 4 + kXcds - 1) / kXcds * kXcds;
    }
    L.nx = (P.nx + 7) / 8;
    L.ny = (P.ny + 7) / 8;
    L.nz = (P.nz + 3) / 4;
    const int layer = L.nx * L.ny;
    const int ring = layer - std::max(L.nx - 2, 0) * std::max(L.ny - 2, 0);
    L.surface = (L.nz >= 2 ? 2 * layer : layer) + std::max(L.nz - 2, 0) * ring;
    L.surface_padded = (L.surface + kXcds - 1) / kXcds * kXcds;
    L.interior = std::max(L.nx - 2, 0) * std::max(L.ny - 2, 0) * std::max(L.nz - 2, 0);
    L.chunk = (L.interior + kXcds - 1) / kXcds;
    return L;
}

// Band coordinate q in 0..3 of an axis of n >= 4 samples: 0, 1, n-2, n-1.
__device__ __forceinline__ int band(int q, int n) { return q < 2 ? q : n - 4 + q; }

__device__ __forceinline__ bool in_band(int i, int n) { return i <= 1 || i >= n - 2; }

// A voxel on an edge line: two of its coordinates in the band.
__device__ __forceinline__ bool on_edge(const int *i, const int *n) {
    const bool x = in_band(i[0], n[0]), y = in_band(i[1], n[1]), z = in_band(i[2], n[2]);
    return (x && y) || (y && z) || (x && z);
}

// Voxel of lane `lane` of edge wave e; false past the line's end.  Lines
// along z take every z at band (x, y); lines along x take x in [2, n-3] at
// band (y, z); lines along y take y in [2, n-3] at band (x, z).
__device__ __forceinline__ bool edge_voxel(const BrickLayout &L, const int *n, int e, int lane, int *i) {
    if (e >= L.edges) return false;
    if (e < 16 * L.zw) {
        const int line = e / L.zw, seg = e - line * L.zw;
        i[0] = band(line & 3, n[0]);
        i[1] = band(line >> 2, n[1]);
        i[2] = seg * 64 + lane;
        return i[2] < n[2];
    }
    e -= 16 * L.zw;
    if (e < 16 * L.xw) {
        const int line = e / L.xw, seg = e - line * L.xw;
        i[1] = band(line & 3, n[1]);
        i[2] = band(line >> 2, n[2]);
        i[0] = 2 + seg * 64 + lane;
        return i[0] < n[0] - 2;
    }
    e -= 16 * L.xw;
    const int line = e / L.yw, seg = e - line * L.yw;
    i[0] = band(line & 3, n[0]);
    i[2] = band(line >> 2, n[2]);
    i[1] = 2 + seg * 64 + lane;
    return i[1] < n[1] - 2;
}

// Brick of workgroup b (counted after the edge section); false for padding
// workgroups.
__device__ __forceinline__ bool brick_of(const BrickLayout &L, int b, int &bx, int &by, int &bz) {
    if (b < L.surface_padded) {
        if (b >= L.surface) return false;
        const int layer = L.nx * L.ny;
        int q;
        if (b < layer) {  // bottom layer
            bz = 0;
            q = b;
        } else if (b >= L.surface - layer) {  // top layer (nz >= 2)
            bz = L.nz - 1;
            q = b - (L.surface - layer);
        } else {  // the ring of a middle layer
            const int ring = layer - max(L.nx - 2, 0) * max(L.ny - 2, 0);
            const int r = b - layer;
            bz = 1 + r / ring;
            q = r - (bz - 1) * ring;
            if (L.nx > 2 && L.ny > 2) {
                // first row, last row, then the two side columns
                if (q < L.nx) {
                    bx = q;
                    by = 0;
                } else if (q < 2 * L.nx) {
                    bx = q - L.nx;
                    by = L.ny - 1;
                } else {
                    const int s = q - 2 * L.nx;
                    by = 1 + (s >> 1);
                    bx = (s & 1) ? L.nx - 1 : 0;
                }
                return true;
            }
        }
        by = q / L.nx;
        bx = q - by * L.nx;
        return true;
    }
    const int i = b - L.surface_padded;
    const int logical = (i % kXcds) * L.chunk + i / kXcds;
    if (logical >= L.interior) return false;
    const int ix = L.nx - 2, iy = L.ny - 2;
    bx = 1 + logical % ix;
    const int rest = logical / ix;
    by = 1 + rest % iy;
    bz = 1 + rest / iy;
    return true;
}

// Voxel bricks of 8 x 8 x 4 per workgroup (4 x 4 x 4 per wave, so a wave's
// voxels see nearly the same rays).
template <typename OUT, bool SIMPLE>
__global__ __launch_bounds__(kBlock) void gather_kernel(Params P, const GatherRay *__restrict__ rays,
                                                        const GatherView *__restrict__ views,
                                                        const int *__restrict__ guard, OUT *__restrict__ out,
                                                        int accumulate, BrickLayout L) {
    if (*guard) return;  // a failed check: the destination stays untouched (uniform)
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int n[3] = {P.nx, P.ny, P.nz};
    Voxel vx;
    // no barriers below: lanes without a voxel return
    if ((int)blockIdx.x < L.edge_groups) {
        if (!edge_voxel(L, n, 4 * (int)blockIdx.x + w, l, vx.i)) return;
    } else {
        int bx, by, bz;
        if (!brick_of(L, (int)blockIdx.x - L.edge_groups, bx, by, bz)) return;  // uniform per workgroup
        vx.i[0] = bx * 8 + (w & 1) * 4 + (l & 3);
        vx.i[1] = by * 8 + (w >> 1) * 4 + ((l >> 2) & 3);
        vx.i[2] = bz * 4 + (l >> 4);
        if (vx.i[0] >= n[0] || vx.i[1] >= n[1] || vx.i[2] >= n[2]) return;
        if (L.edges > 0 && on_edge(vx.i, n)) return;  // gathered by the edge section
    }
    for (int a = 0; a < 3; ++a) {
        vx.c[a] = (double)vx.i[a];
        vx.lo[a] = vx.i[a] <= 1 ? -0.5 : (double)(vx.i[a] - 1);
        vx.hi[a] = vx.i[a] >= n[a] - 2 ? (double)n[a] - 0.5 : (double)(vx.i[a] + 1);
    }
    double total;
    if constexpr (SIMPLE) {
        // axes on which every voxel of the wave is at least two samples
        // from both grid faces (uniform over the wave)
        int mask = 0;
        for (int a = 0; a < 3; ++a) mask |= __all(vx.i[a] >= 2 && vx.i[a] <= n[a] - 3) ? 1 << a : 0;
        switch (mask) {
            case 0: total = gather_voxel<kMasked, 0>(P, rays, views, vx); break;
            case 1: total = gather_voxel<kMasked, 1>(P, rays, views, vx); break;
            case 2: total = gather_voxel<kMasked, 2>(P, rays, views, vx); break;
            case 3: total = gather_voxel<kMasked, 3>(P, rays, views, vx); break;
            case 4: total = gather_voxel<kMasked, 4>(P, rays, views, vx); break;
            case 5: total = gather_voxel<kMasked, 5>(P, rays, views, vx); break;
            case 6: total = gather_voxel<kMasked, 6>(P, rays, views, vx); break;
            default: total = gather_voxel<kMasked, 7>(P, rays, views, vx); break;
        }
    } else {
        total = gather_voxel<kGeneral>(P, rays, views, vx);
    }
    const int64_t voxel = ((int64_t)vx.i[2] * n[1] + vx.i[1]) * n[0] + vx.i[0];
    if (accumulate) total += (double)out[voxel];
    const OUT value = (OUT)total;
    out[voxel] = value;
    if (!isfinite((double)value)) atomicOr(&P.status[0], DPT_STATUS_RANGE);
}

unsigned stride_blocks(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

template <typename OUT>
void launch_voxels(const Params &P, const GatherRay *rays, const GatherView *views, const int *guard, OUT *out,
                   int accumulate, hipStream_t s) {
    const BrickLayout L = brick_layout(P);
    const dim3 grid((unsigned)(L.edge_groups + L.surface_padded + L.chunk * kXcds)), block(kBlock);
    if (P.nx > 1 && P.ny > 1 && P.nz > 1)
        hipLaunchKernelGGL((gather_kernel<OUT, true>), grid, block, 0, s, P, rays, views, guard, out, accumulate, L);
    else
        hipLaunchKernelGGL((gather_kernel<OUT, false>), grid, block, 0, s, P, rays, views, guard, out, accumulate, L);
}

}  // namespace

hipError_t launch_gather(const Params &P, const double *cotangent, GatherRay *rays, GatherView *views, int *guard,
                         void *out, bool out_fp64, bool accumulate, bool checks, const double *panel_meta,
                         const int *panel_modes, int image_cols, hipStream_t s) {
    const int64_t pixels = (int64_t)P.views * P.rows * P.cols;
    hipLaunchKernelGGL(clear_guard_kernel, dim3(1), dim3(1), 0, s, guard);
    if (checks)
        hipLaunchKernelGGL(gather_checks_kernel, dim3(stride_blocks(pixels > P.views ? pixels : P.views)),
                           dim3(kBlock), 0, s, P.poses, P.views, cotangent, pixels, P.status, guard);
    const int image_rows = panel_meta != nullptr ? (int)(P.cols / image_cols) : P.rows;
    hipLaunchKernelGGL(gather_views_kernel, dim3((P.views + 63) / 64), dim3(64), 0, s, P, panel_meta, panel_modes,
                       image_rows, image_cols, views);
    hipLaunchKernelGGL(gather_rays_kernel, dim3(stride_blocks(pixels)), dim3(kBlock), 0, s, P, cotangent, rays);
    if (out_fp64)
        launch_voxels<double>(P, rays, views, guard, static_cast<double *>(out), accumulate ? 1 : 0, s);
    else
        launch_voxels<float>(P, rays, views, guard, static_cast<float *>(out), accumulate ? 1 : 0, s);
    return hipGetLastError();
}

}  // namespace dpt
