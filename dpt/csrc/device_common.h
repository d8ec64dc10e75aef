// Device-side building blocks shared by the projector and the pose-fit
// kernels: launch parameters, FP64 ray setup, FP32 clipping, trilinear
// sampling, slopes, scatter increments and the FP64 pose epilogue.  Every
// function restates the reference operation for operation (citations inline).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <climits>
#include <cstdint>
#include <type_traits>

#include "../../include/dpt_amd.h"

namespace dpt {

constexpr int kBlock = 256;
constexpr int kTile = 16;
constexpr int kXcds = 8;
constexpr int kMaxComponents = 12;
// Component count that selects per-ray source/endpoint derivatives (6 values
// per ray, stored per ray instead of reduced) in the pose epilogue.
constexpr int kRayCoordinates = 7;
// Component count that selects the per-ray 6-DoF local-right tangent (tx, ty,
// tz, rx, ry, rz per ray, stored per ray instead of reduced).
constexpr int kRayTangent = 8;

struct Params {
    int nx, ny, nz;
    int ix, iy;              // row / slab counts of the gather index: (nx, ny) linear, cell counts for cells
    int sx, sy, sz;          // corner strides: 1, nx, nx*ny, or 0 for singleton axes
    float hx, hy, hz;        // n - 0.5 (upper support bound)
    float cx, cy, cz;        // n - 1 (upper clamp)
    int capx, capy, capz;    // max(n - 2, 0)
    double g[9];             // object_to_grid
    double go[3];            // grid origin
    int views, rows, cols, samples;
    int bundle;              // 0: panels, 1: explicit rays
    int vb;                  // views per workgroup (1, 2 or 4): one wave64 per view when > 1
    int tile_w, tile_h;      // pixels per workgroup tile of one view
    int tiles_x, tiles_y, tiles_per_group, total_tiles, chunk;
    int segments;            // FP32 VJP: workgroups per tile, each over its own range of samples
    int base_tiles;          // tiles of one segment (total_tiles / segments)
    int segment_samples;     // samples per segment (a multiple of chunk_samples)
    int slab_fallback;       // FP32 VJP: a chunk that overflows at the shortest length switches the workgroup to slab chunks
    int wt_x, wt_y, wt_per_view;  // wave tiles (8x8 pixels, or 64 rays) per view
    const double *panels;    // views x 12
    const double *ray_src;   // views x pixels x 3
    const double *ray_end;
    const double *poses;     // views x 12
    int *status;             // views
    int brick_cap;           // FP64 LDS brick entries per workgroup (volume adjoint)
    int slab_chunks;         // 1: the FP64 volume adjoint chunks by slabs of space, not by sample index
    int chunk_samples;       // samples per lockstep chunk of the volume adjoint
    int gauss_chunks;        // lockstep chunks per ray of the cell-Gauss volume adjoint
    const float *alpha;      // 2 x samples entries: alpha_k, then k + 0.5 (sample_table_entries)
    int volume_f64;          // 1: FP64 volume destination (increments unflushed, FP64 sampling unrounded)
};

// Per-sample constants, tabulated once on the host so that the sample loop
// reads them with scalar loads of a uniform address instead of computing
// them per lane:
//   alpha_k = RN((k + 0.5) / S) of the pose clip-face shifts
//     (kernels/projection_fp32.py:455-458; IEEE binary32 division, identical
//     to the device's correctly rounded division);
//   c_k = k + 0.5, the sample centre (kernels/projection_fp32.py:116-119;
//     exact in FP32 for k < 2^23).
inline size_t sample_table_entries(int samples) { return 2 * (size_t)samples; }
inline void fill_alpha_table(float *out, int samples) {
    for (int k = 0; k < samples; ++k) {
        out[k] = ((float)k + 0.5f) / (float)samples;
        out[samples + k] = (float)k + 0.5f;
    }
}

// k is uniform across the wave: read the table through the constant address
// space so that the load is a scalar (s_load) one.
__device__ __forceinline__ float sample_alpha(const Params &P, int k) {
    typedef const float __attribute__((address_space(4))) *constant_t;
    return ((constant_t)P.alpha)[k];
}
__device__ __forceinline__ float sample_centre(const Params &P, int k) {
    typedef const float __attribute__((address_space(4))) *constant_t;
    return ((constant_t)P.alpha)[P.samples + k];
}

// ------------------------------------------------------------------ utils --

__device__ __forceinline__ void flag(const Params &P, int view, int bits) {
    atomicOr(&P.status[view], bits);
}

__device__ __forceinline__ bool finite3f(float a, float b, float c) {
    return isfinite(a) && isfinite(b) && isfinite(c);
}

__device__ __forceinline__ bool finite3d(const double *v) {
    return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]);
}

// Workgroup -> rays.  Consecutive workgroup ids land on different XCDs (round
// robin); the remap hands each XCD a contiguous run of the tile walk (view
// group, tile row, tile column) so the tiles an XCD has in flight sample a
// compact part of the volume and share its L2.  A workgroup covers `vb`
// consecutive views: vb = 1 is a 16x16 pixel tile (four 8x8 wave quadrants),
// vb = 2 an 8x16 tile of two views, vb = 4 an 8x8 tile of four views (one wave
// per view).  Each wave owns one 8x8 wave tile of one view; `wslot` indexes its
// pose partials.  Explicit ray bundles use 256 consecutive rays per workgroup.
// Returns false when the whole workgroup lies past the end of the walk.
__device__ __forceinline__ bool lane_ray(const Params &P, int &view, int &row, int &col, int &wslot,
                                         bool &wave_valid) {
    const int b = blockIdx.x;
    int logical = (b % kXcds) * P.chunk + b / kXcds;
    if (logical >= P.total_tiles) return false;
    if (P.segments > 1) logical %= P.base_tiles;  // segment logical / base_tiles
    const int group = logical / P.tiles_per_group;
    const int r = logical - group * P.tiles_per_group;
    const int ty = r / P.tiles_x;
    const int tx = r - ty * P.tiles_x;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    if (P.bundle) {
        view = group;
        row = 0;
        col = tx * kBlock + t;
        const int wx = tx * (kBlock / 64) + w;
        wslot = view * P.wt_per_view + wx;
        wave_valid = wx < P.wt_x;
        return true;
    }
    int vw, qy, qx;
    if (P.vb == 1) {
        vw = 0;
        qy = w >> 1;
        qx = w & 1;
    } else if (P.vb == 2) {
        vw = w >> 1;
        qy = 0;
        qx = w & 1;
    } else {
        vw = w;
        qy = 0;
        qx = 0;
    }
    view = group * P.vb + vw;
    const int wy = ty * (P.tile_h >> 3) + qy, wx = tx * (P.tile_w >> 3) + qx;
    row = wy * 8 + (l >> 3);
    col = wx * 8 + (l & 7);
    wslot = view * P.wt_per_view + wy * P.wt_x + wx;
    wave_valid = view < P.views && wy < P.wt_y && wx < P.wt_x;
    return true;
}

__device__ __forceinline__ bool ray_in_range(const Params &P, int view, int row, int col) {
    return view < P.views && row < P.rows && col < P.cols;
}

// ------------------------------------------------------------- ray setup --
// kernels/projection.py:79-102 in FP64; same operation order as the oracle.

struct World {
    double wsr[3], wd[3], os[3], od[3];
};

__device__ __forceinline__ void ray_world(const Params &P, const double *pose, int view, int row,
                                          int col, World &w) {
    double src[3], end[3];
    if (!P.bundle) {
        const double *pn = P.panels + 12 * (size_t)view;
        for (int i = 0; i < 3; ++i) {
            src[i] = pn[i];
            end[i] = pn[3 + i] + (double)col * pn[6 + i] + (double)row * pn[9 + i];
        }
    } else {
        const size_t ray = (size_t)view * P.cols + col;
        for (int i = 0; i < 3; ++i) {
            src[i] = P.ray_src[3 * ray + i];
            end[i] = P.ray_end[3 * ray + i];
        }
    }
    for (int i = 0; i < 3; ++i) {
        w.wsr[i] = src[i] - pose[9 + i];
        w.wd[i] = end[i] - src[i];
    }
    for (int i = 0; i < 3; ++i) {
        w.os[i] = pose[i] * w.wsr[0] + pose[3 + i] * w.wsr[1] + pose[6 + i] * w.wsr[2];
        w.od[i] = pose[i] * w.wd[0] + pose[3 + i] * w.wd[1] + pose[6 + i] * w.wd[2];
    }
}

struct Ray {
    double o[3], d[3], len;
    bool valid;
};

__device__ __forceinline__ void ray_grid(const Params &P, const World &w, Ray &r) {
    double rel[3];
    for (int i = 0; i < 3; ++i) rel[i] = w.os[i] - P.go[i];
    for (int i = 0; i < 3; ++i) {
        r.o[i] = P.g[3 * i] * rel[0] + P.g[3 * i + 1] * rel[1] + P.g[3 * i + 2] * rel[2];
        r.d[i] = P.g[3 * i] * w.od[0] + P.g[3 * i + 1] * w.od[1] + P.g[3 * i + 2] * w.od[2];
    }
    r.len = sqrt(w.wd[0] * w.wd[0] + w.wd[1] * w.wd[1] + w.wd[2] * w.wd[2]);
    // kernels/projection.py:110-121 (valid_ray)
    r.valid = finite3d(r.o) && finite3d(r.d) && finite3d(w.os) && finite3d(w.od) &&
              finite3d(w.wsr) && finite3d(w.wd) && isfinite(r.len) && r.len > 0.0;
}

// --------------------------------------------------------- FP64 helpers --

struct Interval64 {
    double lower, upper;
    int lo_axis, up_axis;
    double lo_o, lo_d, up_o, up_d;
};

// kernels/projection.py:147-184
__device__ __forceinline__ Interval64 interval64(const Params &P, const double *o, const double *d) {
    Interval64 iv;
    iv.lower = 0.0;
    iv.upper = 1.0;
    iv.lo_axis = iv.up_axis = -1;
    iv.lo_o = iv.lo_d = iv.up_o = iv.up_d = 0.0;
    const double highs[3] = {(double)P.nx - 0.5, (double)P.ny - 0.5, (double)P.nz - 0.5};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double low = -0.5, high = highs[a], v = d[a];
        if (v == 0.0) {
            if (o[a] < low || o[a] > high) iv.upper = -1.0;
        } else {
            double first = (low - o[a]) / v, last = (high - o[a]) / v;
            if (first > last) {
                const double t = first;
                first = last;
                last = t;
            }
            if (first > iv.lower) {
                iv.lower = first;
                iv.lo_axis = a;
                iv.lo_o = -1.0 / v;
                iv.lo_d = -first / v;
            }
            if (last < iv.upper) {
                iv.upper = last;
                iv.up_axis = a;
                iv.up_o = -1.0 / v;
                iv.up_d = -last / v;
            }
        }
    }
    return iv;
}

// kernels/numerics.py:11-56: product through a mantissa/exponent split
__device__ __forceinline__ double product4(double a, double b, double c, double d) {
    const double f[4] = {a, b, c, d};
    double ordinary = a * b * c * d;
    bool invalid = false, zero = false;
    for (int i = 0; i < 4; ++i) {
        invalid |= !isfinite(f[i]);
        zero |= f[i] == 0.0;
    }
    if (invalid) return ordinary;
    if (zero) return 0.0;
    double m = 1.0;
    int e = 0;
    for (int i = 0; i < 4; ++i) {
        int ei;
        const double mi = frexp(f[i], &ei);
        m = (i == 0) ? mi : m * mi;
        e += ei;
    }
    return ldexp(m, e);
}

// kernels/numerics.py:11-56 for three factors
__device__ __forceinline__ double product3(double a, double b, double c) {
    const double f[3] = {a, b, c};
    double ordinary = a * b * c;
    bool invalid = false, zero = false;
    for (int i = 0; i < 3; ++i) {
        invalid |= !isfinite(f[i]);
        zero |= f[i] == 0.0;
    }
    if (invalid) return ordinary;
    if (zero) return 0.0;
    double m = 1.0;
    int e = 0;
    for (int i = 0; i < 3; ++i) {
        int ei;
        const double mi = frexp(f[i], &ei);
        m = (i == 0) ? mi : m * mi;
        e += ei;
    }
    return ldexp(m, e);
}


// kernels/batched_geometry.py:71-88 + kernels/geometry.py:55-74: finite
// entries, orthonormal to 1e-10 and a proper rotation.
__device__ __forceinline__ bool pose_is_rigid(const double *R) {
    bool valid = isfinite(R[9]) && isfinite(R[10]) && isfinite(R[11]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            if (!isfinite(R[3 * i + j])) valid = false;
            // (R^T R)[i][j] = sum_k R[k][i] R[k][j]
            const double prod = R[i] * R[j] + R[3 + i] * R[3 + j] + R[6 + i] * R[6 + j];
            const double expected = (i == j) ? 1.0 : 0.0;
            if (!(fabs(prod - expected) <= 1e-10)) valid = false;
        }
    const double det = R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                       R[2] * (R[3] * R[7] - R[4] * R[6]);
    if (det < 0.0) valid = false;
    return valid;
}

// --------------------------------------------------------- FP32 sampling --

struct Interval32 {
    float lower, upper;
    int lo_axis, up_axis;    // active clip face (-1: none)
    float lo_o, lo_d, up_o, up_d;
};

// kernels/projection_fp32.py:77-109; only the single active face derivative
// is non-zero, so the interval keeps its axis and two scalars per end.
__device__ __forceinline__ Interval32 interval32(const Params &P, const float *o, const float *d) {
    Interval32 iv;
    iv.lower = 0.0f;
    iv.upper = 1.0f;
    iv.lo_axis = iv.up_axis = -1;
    iv.lo_o = iv.lo_d = iv.up_o = iv.up_d = 0.0f;
    const float highs[3] = {P.hx, P.hy, P.hz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float low = -0.5f, high = highs[a], v = d[a];
        if (v == 0.0f) {
            if (o[a] < low || o[a] > high) iv.upper = -1.0f;
        } else {
            float first = (low - o[a]) / v, last = (high - o[a]) / v;
            if (first > last) {
                const float t = first;
                first = last;
                last = t;
            }
            if (first > iv.lower) {
                iv.lower = first;
                iv.lo_axis = a;
                iv.lo_o = -1.0f / v;
                iv.lo_d = -first / v;
            }
            if (last < iv.upper) {
                iv.upper = last;
                iv.up_axis = a;
                iv.up_o = -1.0f / v;
                iv.up_d = -last / v;
            }
        }
    }
    return iv;
}

template <typename T>
__device__ __forceinline__ float load_value(const T *f, int i);
template <>
__device__ __forceinline__ float load_value<float>(const float *f, int i) {
    return f[i];
}
template <>
__device__ __forceinline__ float load_value<__hip_bfloat16>(const __hip_bfloat16 *f, int i) {
    const uint16_t bits = reinterpret_cast<const uint16_t *>(f)[i];
    return __uint_as_float((uint32_t)bits << 16);
}
template <>
__device__ __forceinline__ float load_value<__half>(const __half *f, int i) {
    return __half2float(f[i]);
}

struct Cell32 {
    int base;                // linear index of the lower corner
    int lx, ly, lz;          // lower corner
    float fx, fy, fz;        // fractions
    float slx, sly, slz;     // slopes (0/1)
};

// Maximum ny * nz of a grid: gather indices use 24-bit multiplies
// ((lz * iy + ly) * ix + lx with both products' operands below 2^24).
constexpr long long kMaxSliceProduct = 1LL << 24;

// Support test of one FP32 sample point (kernels/projection_fp32.py:64-74):
// NaN and infinities fail the comparisons, as finite_point does.
__device__ __forceinline__ bool inside32(const Params &P, float p0, float p1, float p2) {
    return p0 >= -0.5f && p1 >= -0.5f && p2 >= -0.5f && p0 <= P.hx && p1 <= P.hy && p2 <= P.hz;
}

// kernels/projection_fp32.py:130-140 for a point inside the support.  For
// the finite coordinates that reach the clamp, med3 equals
// min(max(p, 0), n - 1), and min(trunc(v), n - 2) equals the reference's
// integer min of floor(v) (v >= 0).
__device__ __forceinline__ void cell32_inside(const Params &P, float p0, float p1, float p2, Cell32 &c) {
    const float v0 = __builtin_amdgcn_fmed3f(p0, 0.0f, P.cx);
    const float v1 = __builtin_amdgcn_fmed3f(p1, 0.0f, P.cy);
    const float v2 = __builtin_amdgcn_fmed3f(p2, 0.0f, P.cz);
    c.lx = min((int)v0, P.capx);
    c.ly = min((int)v1, P.capy);
    c.lz = min((int)v2, P.capz);
    // v - lower is exact (Sterbenz)
    c.fx = v0 - (float)c.lx;
    c.fy = v1 - (float)c.ly;
    c.fz = v2 - (float)c.lz;
    c.slx = (p0 >= 0.0f && p0 < P.cx) ? 1.0f : 0.0f;
    c.sly = (p1 >= 0.0f && p1 < P.cy) ? 1.0f : 0.0f;
    c.slz = (p2 >= 0.0f && p2 < P.cz) ? 1.0f : 0.0f;
    // gather index: the voxel of the lower corner (linear layouts) or the
    // cell (cell layout), by the row / slab counts of the launch
    c.base = (int)(__umul24(__umul24((unsigned)c.lz, (unsigned)P.iy) + (unsigned)c.ly, (unsigned)P.ix) +
                   (unsigned)c.lx);
}

// Returns false outside the support.
__device__ __forceinline__ bool cell32(const Params &P, float p0, float p1, float p2, Cell32 &c) {
    if (!inside32(P, p0, p1, p2)) return false;
    cell32_inside(P, p0, p1, p2, c);
    return true;
}

// corner k value offsets: bit0 x, bit1 y, bit2 z (kernels/projection.py:217-222;
// lower <= n-2 so only singleton axes clamp, which a zero stride reproduces).
__device__ __forceinline__ int corner_offset(const Params &P, int k) {
    return ((k & 1) ? P.sx : 0) + ((k & 2) ? P.sy : 0) + ((k & 4) ? P.sz : 0);
}

// corner weights (f0*f1)*f2 (kernels/projection_fp32.py:143-154)
__device__ __forceinline__ void weights32(const Cell32 &c, float *w) {
    const float ax = 1.0f - c.fx, ay = 1.0f - c.fy, az = 1.0f - c.fz;
    const float xy[4] = {ax * ay, c.fx * ay, ax * c.fy, c.fx * c.fy};
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = xy[k & 3] * ((k & 4) ? c.fz : az);
}

// Eight corner values.  With PAIR (FP32 storage, nx > 1) each x-adjacent pair
// (lower x, lower x + 1 <= nx - 1) is one dword-aligned 8-byte load: four
// loads per sample instead of eight, never past the end of an x row.
struct __attribute__((packed, aligned(4))) FloatPair {
    float a, b;
};
struct __attribute__((packed, aligned(4))) UintPair {
    uint32_t a, b;
};

template <typename T>
__device__ __forceinline__ float half_bits_to_float(uint32_t bits);
template <>
__device__ __forceinline__ float half_bits_to_float<__hip_bfloat16>(uint32_t bits) {
    return __uint_as_float(bits << 16);
}
template <>
__device__ __forceinline__ float half_bits_to_float<__half>(uint32_t bits) {
    return __half2float(__ushort_as_half((unsigned short)bits));
}

// Field buffers always live in global memory.  Loads go through pointers of
// this type so that a field pointer read from a device table (not known to the
// compiler as global) still compiles to global_load rather than flat_load.
template <typename T>
using global_const = const T __attribute__((address_space(1)));

template <typename T>
__device__ __forceinline__ float load_global(global_const<T> *f, int i);
template <>
__device__ __forceinline__ float load_global<float>(global_const<float> *f, int i) {
    return f[i];
}
template <>
__device__ __forceinline__ float load_global<__hip_bfloat16>(global_const<__hip_bfloat16> *f, int i) {
    const uint16_t bits = reinterpret_cast<global_const<uint16_t> *>(f)[i];
    return __uint_as_float((uint32_t)bits << 16);
}
template <>
__device__ __forceinline__ float load_global<__half>(global_const<__half> *f, int i) {
    const uint16_t bits = reinterpret_cast<global_const<uint16_t> *>(f)[i];
    return __half2float(__ushort_as_half(bits));
}

// Gather layouts of the FP32-sampling kernels.
constexpr int kLayoutScalar = 0;  // linear field, eight scalar loads (nx == 1)
constexpr int kLayoutPair = 1;    // linear field, x-adjacent pairs (nx > 1)
constexpr int kLayoutCells = 2;   // cell layout (DPT_STORE_CELLS): one cell per gather

struct __attribute__((aligned(16))) Quad {
    uint32_t x, y, z, w;
};

template <typename T, int LAYOUT>
__device__ __forceinline__ void load8(const Params &P, const T *field_any, int base, float *v) {
    global_const<T> *field = (global_const<T> *)field_any;
    if constexpr (LAYOUT == kLayoutCells && std::is_same<T, float>::value) {
        // 32-byte cell: corners 0-3 then 4-7, two 16-byte loads
        global_const<Quad> *q = reinterpret_cast<global_const<Quad> *>(field) + 2 * (size_t)base;
        const Quad a = {q[0].x, q[0].y, q[0].z, q[0].w};
        const Quad b = {q[1].x, q[1].y, q[1].z, q[1].w};
        v[0] = __uint_as_float(a.x);
        v[1] = __uint_as_float(a.y);
        v[2] = __uint_as_float(a.z);
        v[3] = __uint_as_float(a.w);
        v[4] = __uint_as_float(b.x);
        v[5] = __uint_as_float(b.y);
        v[6] = __uint_as_float(b.z);
        v[7] = __uint_as_float(b.w);
    } else if constexpr (LAYOUT == kLayoutCells) {
        // 16-byte cell of 16-bit values, corner k in half k
        global_const<Quad> *q = reinterpret_cast<global_const<Quad> *>(field) + (size_t)base;
        const uint32_t w[4] = {q->x, q->y, q->z, q->w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[2 * j] = half_bits_to_float<T>(w[j] & 0xFFFFu);
            v[2 * j + 1] = half_bits_to_float<T>(w[j] >> 16);
        }
    } else if constexpr (LAYOUT == kLayoutPair && std::is_same<T, float>::value) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            global_const<FloatPair> *q = reinterpret_cast<global_const<FloatPair> *>(
                field + base + ((j & 1) ? P.sy : 0) + ((j & 2) ? P.sz : 0));
            v[2 * j] = q->a;
            v[2 * j + 1] = q->b;
        }
    } else if constexpr (LAYOUT == kLayoutPair) {
        // 16-bit storage: the x-adjacent pair (i, i+1) lies in the dword-aligned
        // 8 bytes at dword i/2; one dwordx2 load and a funnel shift extract it.
        // Reads at most 4 bytes past element i+1: half-storage buffers carry
        // DPT_HALF_PAD_BYTES of padding.
        global_const<uint32_t> *words = reinterpret_cast<global_const<uint32_t> *>(field);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = base + ((j & 1) ? P.sy : 0) + ((j & 2) ? P.sz : 0);
            global_const<UintPair> *q = reinterpret_cast<global_const<UintPair> *>(words + (i >> 1));
            const uint32_t pair = __builtin_amdgcn_alignbit(q->b, q->a, (i & 1) * 16);
            v[2 * j] = half_bits_to_float<T>(pair & 0xFFFFu);
            v[2 * j + 1] = half_bits_to_float<T>(pair >> 16);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = load_global<T>(field, base + corner_offset(P, k));
    }
}

// sum_k fma(w_k, v_k, .) in corner order (kernels/projection_fp32.py:165-168)
__device__ __forceinline__ float interpolate32(const float *w, const float *v) {
    float value = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) value = fmaf(w[k], v[k], value);
    return value;
}

// kernels/projection_fp32.py:196-208 (differences before interpolation)
__device__ __forceinline__ void slopes32(const Cell32 &c, const float *v, float *g) {
    const float fx = c.fx, fy = c.fy, fz = c.fz;
    const float ax = 1.0f - fx, ay = 1.0f - fy, az = 1.0f - fz;
    const float x0 = fmaf(fy, v[3] - v[2], ay * (v[1] - v[0]));
    const float x1 = fmaf(fy, v[7] - v[6], ay * (v[5] - v[4]));
    g[0] = c.slx * fmaf(fz, x1, az * x0);
    const float y0 = fmaf(fx, v[3] - v[1], ax * (v[2] - v[0]));
    const float y1 = fmaf(fx, v[7] - v[5], ax * (v[6] - v[4]));
    g[1] = c.sly * fmaf(fz, y1, az * y0);
    const float z0 = fmaf(fx, v[5] - v[1], ax * (v[4] - v[0]));
    const float z1 = fmaf(fx, v[7] - v[3], ax * (v[6] - v[2]));
    g[2] = c.slz * fmaf(fy, z1, ay * z0);
}

// kernels/projection_fp32.py:240-252: RN(factor*weight), |x| < 2^-126 -> signed 0
// One FP64-sampled corner contribution as the volume destination receives
// it: rounded once to FP32 for an FP32 destination, unrounded for an FP64
// one (kernels/projection.py:407-430, get_scatter_realised).
__device__ __forceinline__ double destination_increment(const Params &P, double contribution) {
    return P.volume_f64 ? contribution : (double)(float)contribution;
}

__device__ __forceinline__ float scatter_increment(float factor, float weight) {
    const float x = factor * weight;
    return fabsf(x) < 0x1p-126f ? copysignf(0.0f, x) : x;
}

// Compensated (Kahan) accumulator, used only by half-storage kernels.
struct Kahan {
    float s, c;
    __device__ __forceinline__ void init() { s = c = 0.0f; }
    __device__ __forceinline__ void add(float x) {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// Along-ray accumulator.  Plain: one FP32 running sum, the reference's
// arithmetic.  Compensated (half-storage forms): blocked compensated
// summation -- an FP32 partial sum over each block of samples, folded into a
// Kahan-compensated total at block boundaries -- which keeps the error of a
// 1024-sample ray near that of full Kahan at ~1 extra add per sample.
constexpr int kFoldBlock = 32;

template <bool COMPENSATED>
struct Acc {
    float s;
    __device__ __forceinline__ void init() { s = 0.0f; }
    __device__ __forceinline__ void add(float x) { s += x; }
    __device__ __forceinline__ void fma_add(float a, float b) { s = fmaf(a, b, s); }
    __device__ __forceinline__ void fold() {}
    __device__ __forceinline__ float value() const { return s; }
};
template <>
struct Acc<true> {
    float blk;
    Kahan k;
    __device__ __forceinline__ void init() {
        blk = 0.0f;
        k.init();
    }
    __device__ __forceinline__ void add(float x) { blk += x; }
    __device__ __forceinline__ void fma_add(float a, float b) { blk = fmaf(a, b, blk); }
    __device__ __forceinline__ void fold() {
        k.add(blk);
        blk = 0.0f;
    }
    __device__ __forceinline__ float value() const {
        Kahan t = k;
        t.add(blk);
        return t.s;
    }
};

// Position of sample k and its interpolation cell; c.base = -1 outside the
// support (kernels/projection_fp32.py:116-127, 130-140).  CHECK = false skips
// the support test for a span of samples that span_inside has cleared.
template <bool CHECK = true>
__device__ __forceinline__ void sample_at(const Params &P, const float *o, const float *d, float lower, float step,
                                          int k, float &pos, Cell32 &c) {
    pos = fmaf(sample_centre(P, k), step, lower);
    const float p0 = fmaf(pos, d[0], o[0]);
    const float p1 = fmaf(pos, d[1], o[1]);
    const float p2 = fmaf(pos, d[2], o[2]);
    if constexpr (CHECK) {
        if (!cell32(P, p0, p1, p2, c)) c.base = -1;
    } else {
        cell32_inside(P, p0, p1, p2, c);
    }
}

// Every sample point of [k0, k1) lies in the support iff the first and the
// last do: pos_k = RN(c_k * step + lower) is non-decreasing in k (step > 0)
// and each coordinate RN(pos * d_a + o_a) is monotone in pos, so the samples
// inside the (convex) support along each axis form a run of consecutive k.
__device__ __forceinline__ bool span_inside(const Params &P, const float *o, const float *d, float lower, float step,
                                            int k0, int k1) {
    const float a = fmaf(sample_centre(P, k0), step, lower), b = fmaf(sample_centre(P, k1 - 1), step, lower);
    return inside32(P, fmaf(a, d[0], o[0]), fmaf(a, d[1], o[1]), fmaf(a, d[2], o[2])) &&
           inside32(P, fmaf(b, d[0], o[0]), fmaf(b, d[1], o[1]), fmaf(b, d[2], o[2]));
}

// The reference flags any non-finite sample point (projection_fp32.py:366-367).
// Sample positions are monotone in the sample index, so every point of samples
// [k0, k1) is finite iff the first and the last are.
__device__ __forceinline__ bool finite_span(const float *o, const float *d, float lower, float step, int k0, int k1) {
    const float a = fmaf((float)k0 + 0.5f, step, lower), b = fmaf((float)(k1 - 1) + 0.5f, step, lower);
    return finite3f(fmaf(a, d[0], o[0]), fmaf(a, d[1], o[1]), fmaf(a, d[2], o[2])) &&
           finite3f(fmaf(b, d[0], o[0]), fmaf(b, d[1], o[1]), fmaf(b, d[2], o[2]));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// FP64 pose epilogue: kernels/projection_fp32.py:562-602 / projection.py:1156-1217
template <int C>
__device__ __forceinline__ void pose_epilogue(const Params &P, const double *pose, const World &w,
                                              double factor, const double *og, const double *dg,
                                              double *out, double seed = 0.0, double integral = 0.0,
                                              bool fp64_form = false) {
    double sa[3], da[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double s = P.g[i] * og[0] + P.g[3 + i] * og[1] + P.g[6 + i] * og[2];
        const double t = P.g[i] * dg[0] + P.g[3 + i] * dg[1] + P.g[6 + i] * dg[2];
        sa[i] = factor * s;
        da[i] = factor * t;
    }
    if constexpr (C == kRayCoordinates) {
        // kernels/projection_fp32.py:570-583 / kernels/projection.py:1187-1200:
        // L = |e - s| I(s, e - s), so the segment norm adds seed * I * (e - s)/|e - s|
        // to the endpoint; the FP64 path forms that product range-safely.
        double sw[3], dw[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            sw[i] = pose[3 * i] * sa[0] + pose[3 * i + 1] * sa[1] + pose[3 * i + 2] * sa[2];
            dw[i] = pose[3 * i] * da[0] + pose[3 * i + 1] * da[1] + pose[3 * i + 2] * da[2];
        }
        const double length = sqrt(w.wd[0] * w.wd[0] + w.wd[1] * w.wd[1] + w.wd[2] * w.wd[2]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            dw[a] += fp64_form ? product3(seed, integral, w.wd[a] / length) : seed * integral * w.wd[a] / length;
            out[a] = sw[a] - dw[a];
            out[a + 3] = dw[a];
        }
    } else if constexpr (C == 12) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) out[3 * r + c] = w.wsr[r] * sa[c] + w.wd[r] * da[c];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            out[9 + a] = -(pose[3 * a] * sa[0] + pose[3 * a + 1] * sa[1] + pose[3 * a + 2] * sa[2]);
    } else {
        const double *os = w.os, *od = w.od;
        out[0] = -sa[0];
        out[1] = -sa[1];
        out[2] = -sa[2];
        out[3] = (sa[1] * os[2] - sa[2] * os[1]) + (da[1] * od[2] - da[2] * od[1]);
        out[4] = (sa[2] * os[0] - sa[0] * os[2]) + (da[2] * od[0] - da[0] * od[2]);
        out[5] = (sa[0] * os[1] - sa[1] * os[0]) + (da[0] * od[1] - da[1] * od[0]);
    }
}


}  // namespace dpt
