// MI355X (gfx950) cone-beam projector: forward ray march, volume VJP and
// 6-DoF pose VJP, FP32 and FP64 sampling, FP32/BF16/FP16 field storage.
//
// Semantics follow the reference kernels operation for operation:
//   FP32 sampling  kernels/projection_fp32.py:59-617
//   FP64 sampling  kernels/projection.py:79-296, 396-430, 802-1237
//   ray setup      kernels/projection.py:79-121 (FP64, once per ray)
//   pose algebra   kernels/geometry.py:13-74
// The translation unit is compiled with -ffp-contract=off; every FMA below is
// written explicitly (fmaf/fma).  Where the reference leaves a product-sum
// unmarked and its code generator (Warp, fuse_fp=True) fuses it, the fused form
// is written out, so the per-ray results are a fixed function of the inputs.
//
// Execution model (not the reference's): one launch covers all views.  A
// workgroup of 256 lanes owns one 16x16 detector tile of one view, or 8x8
// tiles of four consecutive views (one wave64 each); workgroup ids are
// remapped so each XCD walks a contiguous range of tiles.  Corners are
// gathered from the linear field (x-adjacent pairs) or from a cell-layout copy
// (8 corners per cell, one or two 16-byte loads).  Pose cotangents are
// reduced per ray in FP64, across the wave with shuffles, then across tiles in
// a fixed-order FP64 pass.  Volume cotangents: a ray whose factor
// f = seed * length * step is finite with |f| >= 2^-102 sums its corner
// weights per cell in FP32 registers and scales each sum by f in FP64 when the
// cell is flushed into an FP64 LDS brick of the cells the workgroup touches
// (the brick is drained into an FP64 accumulator with one device atomic per
// touched grid sample); other rays add the reference's per-increment
// RN(f * w), flushed to zero below 2^-126, with device atomics.  A final pass
// rounds the accumulator (plus the previous gradient when accumulating) once
// to FP32, or keeps FP64 for an FP64 destination.

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <algorithm>
#include <climits>
#include <mutex>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/dpt_amd.h"

#include "device_common.h"
#include "gather.h"

namespace dpt {

// PIPE (FP32 storage; forward and pose-only VJP): the samples of a ray in one
// software pipeline, the corner loads of sample k + kPipeDepth issued before sample k
// is summed.  Small launches (one wave per SIMD or fewer) expose a memory
// latency per sample; this hides it without splitting the ray, so each pose
// sum stays the reference's one sequential FP32 sum (forward: the same
// sequential sum of interpolated values).  Eight slots of corner
// values, fractions and positions take 220 VGPRs (two waves per SIMD); on a
// one-view 512^2 launch depth 8 ran the pose VJP in 0.61 ms (central-window
// cotangent) and 0.84 ms (every ray), depth 4 in 0.86 / 1.31 ms, and
// recomputing the cell at the sum instead of keeping it was slower.
constexpr int kPipeDepth = 8;

// ------------------------------------------------------- FP32 forward ------
// kernels/projection_fp32.py:334-374

template <typename T, bool COMPENSATED, int LAYOUT, bool PIPE = false>
__global__ __launch_bounds__(kBlock) void forward32_kernel(Params P, const T *__restrict__ field,
                                                           float *__restrict__ out) {
    int view, row, col, wslot;
    bool wave_valid;
    if (!lane_ray(P, view, row, col, wslot, wave_valid)) return;
    if (!ray_in_range(P, view, row, col)) return;
    const size_t ray = ((size_t)view * P.rows + row) * P.cols + col;
    const double *pose = P.poses + 12 * (size_t)view;
    World w;
    ray_world(P, pose, view, row, col, w);
    Ray r;
    ray_grid(P, w, r);
    float o[3] = {(float)r.o[0], (float)r.o[1], (float)r.o[2]};
    float d[3] = {(float)r.d[0], (float)r.d[1], (float)r.d[2]};
    const float len = (float)r.len;
    if (!r.valid || !finite3f(o[0], o[1], o[2]) || !finite3f(d[0], d[1], d[2]) || !isfinite(len) ||
        len <= 0.0f) {
        flag(P, view, DPT_STATUS_RANGE);
        out[ray] = 0.0f;
        return;
    }
    const Interval32 iv = interval32(P, o, d);
    Acc<COMPENSATED> total;
    total.init();
    float result = 0.0f;
    if (iv.upper > iv.lower) {
        const float step = (iv.upper - iv.lower) / (float)P.samples;
        bool bad = false;
        // Lockstep sample walk with the current cell's corners cached in
        // registers: only lanes whose cell changed issue the (masked) loads.
        bad = !finite_span(o, d, iv.lower, step, 0, P.samples);
        float v[8];
        int cur = -1;
        auto walk = [&](auto check) {
            constexpr bool CHECK = decltype(check)::value;
            for (int k = 0; k < P.samples; ++k) {
                if constexpr (COMPENSATED) {
                    if (k > 0 && (k % kFoldBlock) == 0) total.fold();
                }
                float pos;
                Cell32 c;
                sample_at<CHECK>(P, o, d, iv.lower, step, k, pos, c);
                if constexpr (CHECK) {
                    if (c.base < 0) continue;
                }
                if (c.base != cur) {
                    load8<T, LAYOUT>(P, field, c.base, v);
                    cur = c.base;
                }
                float wt[8];
                weights32(c, wt);
                total.add(interpolate32(wt, v));
            }
        };
        static_assert(!PIPE || !COMPENSATED, "the pipeline is an FP32-storage loop");
        if constexpr (PIPE) {
            // slot j: corner values, fractions and support flag of the sample
            // issued kPipeDepth samples back; an outside sample adds -0, the
            // identity of FP32 addition (the reference skips it)
            float ring[kPipeDepth][8], rf[kPipeDepth][3], cur[8];
            bool rin[kPipeDepth], fresh[kPipeDepth];
            int prev = -1;
            auto issue = [&](int j, int k) {
                const float pos = fmaf(sample_centre(P, k), step, iv.lower);
                const float p0 = fmaf(pos, d[0], o[0]);
                const float p1 = fmaf(pos, d[1], o[1]);
                const float p2 = fmaf(pos, d[2], o[2]);
                Cell32 c;
                cell32_inside(P, p0, p1, p2, c);  // clamped: the load stays in the field
                rin[j] = inside32(P, p0, p1, p2);
                rf[j][0] = c.fx;
                rf[j][1] = c.fy;
                rf[j][2] = c.fz;
                // only lanes whose cell changed since the previous sample load
                fresh[j] = c.base != prev;
                prev = c.base;
                if (fresh[j]) load8<T, LAYOUT>(P, field, c.base, ring[j]);
            };
            auto consume = [&](int j) {
#pragma unroll
                for (int q = 0; q < 8; ++q) cur[q] = fresh[j] ? ring[j][q] : cur[q];
                Cell32 c;
                c.fx = rf[j][0];
                c.fy = rf[j][1];
                c.fz = rf[j][2];
                float wt[8];
                weights32(c, wt);
                const float value = interpolate32(wt, cur);
                total.add(rin[j] ? value : -0.0f);
            };
#pragma unroll
            for (int q = 0; q < 8; ++q) cur[q] = 0.0f;
            const int last = P.samples;
#pragma unroll
            for (int j = 0; j < kPipeDepth; ++j) issue(j, min(j, last - 1));
            int k0 = 0;
            for (; k0 + kPipeDepth <= last; k0 += kPipeDepth) {
#pragma unroll
                for (int j = 0; j < kPipeDepth; ++j) {
                    consume(j);
                    issue(j, min(k0 + j + kPipeDepth, last - 1));
                }
            }
#pragma unroll
            for (int j = 0; j < kPipeDepth; ++j)
                if (k0 + j < last) consume(j);  // uniform
        } else if (__builtin_amdgcn_ballot_w64(!span_inside(P, o, d, iv.lower, step, 0, P.samples)) == 0) {
            // per-sample support tests only for waves with a ray whose end
            // samples leave the support
            walk(std::false_type{});
        } else {
            walk(std::true_type{});
        }
        if (bad) flag(P, view, DPT_STATUS_RANGE);
        result = total.value() * (len * step);
    }
    out[ray] = result;
    if (!isfinite(result)) flag(P, view, DPT_STATUS_RANGE);
}

// ----------------------------------------------------- pose reduction -----

// Sum C components over the 64 lanes of the wave (fixed shuffle tree); lane 0
// writes the wave tile's partials.
template <int C>
__device__ __forceinline__ void wave_store(const double *g, double *partial, bool valid) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double s = wave_sum(g[c]);
        if (valid && (threadIdx.x & 63) == 0) partial[c] = s;
    }
}

// Store the epilogue's result: per ray for ray coordinates and per-ray
// tangents (6 values at partials + 6 * ray, non-finite flagged), else the
// wave tile's sum.
template <int C>
__device__ __forceinline__ void store_gradient(const Params &P, const double *g, double *partials, int wslot,
                                               bool wave_valid, int view, int row, int col) {
    if constexpr (C == kRayCoordinates || C == kRayTangent) {
        if (!ray_in_range(P, view, row, col)) return;
        const size_t ray = ((size_t)view * P.rows + row) * P.cols + col;
        bool bad = false;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            partials[6 * ray + c] = g[c];
            bad |= !isfinite(g[c]);
        }
        if (bad) flag(P, view, DPT_STATUS_RANGE);
    } else {
        wave_store<C>(g, partials + (size_t)wslot * C, wave_valid);
    }
}

// Register cell accumulator of the volume cotangent (the reference's Vec8
// cell sums, kernels/projection_fp32.py:261-269): a lane sums f * w over its
// consecutive samples in one interpolation cell, one FP32 fused multiply-add
// per corner and sample, and flushes the cell once when it leaves it.
struct CellSums {
    int x, y, z;   // lower corner; x < 0 when empty
    float s[8];    // FP32 sums of f * w of the cell's samples
    __device__ __forceinline__ void reset() {
        x = -1;
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] = 0.0f;
    }
};

// Brick = axis-aligned box of grid samples that one workgroup's rays touch
// during one chunk of samples, privatised in LDS.
struct Brick {
    int x0, y0, z0;      // first sample of the box
    int bx, by, bz;      // extents
    int lsx, lsy, lsz;   // local corner strides (0 on singleton grid axes)
    int volume;
    int origin;          // (z0 * by + y0) * bx + x0: local index = (z * by + y) * bx + x - origin
};

// Add a cell's sums to the LDS brick: eight FP64 LDS atomics (every corner;
// exact-boundary fractions aside, a visited cell's sums are non-zero).
__device__ __forceinline__ void flush_lds(const Brick &B, const CellSums &cs, double *brick) {
    // z * by + y < nz * (ny + 1) <= 2^24 and the index below 2^31 (grid check,
    // stride padding only where it keeps both)
    const int base =
        (int)(__umul24(__umul24((unsigned)cs.z, (unsigned)B.by) + (unsigned)cs.y, (unsigned)B.bx) + (unsigned)cs.x) -
        B.origin;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int off = ((k & 1) ? B.lsx : 0) + ((k & 2) ? B.lsy : 0) + ((k & 4) ? B.lsz : 0);
        atomicAdd(&brick[base + off], (double)cs.s[k]);
    }
}

// Add a cell's non-zero sums to the FP64 device accumulator (chunks whose
// brick does not fit LDS).
__device__ __forceinline__ void flush_device(const Params &P, const CellSums &cs, double *acc) {
    const int base = (cs.z * P.ny + cs.y) * P.nx + cs.x;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (cs.s[k] != 0.0f) unsafeAtomicAdd(&acc[base + corner_offset(P, k)], (double)cs.s[k]);
}

// Chunks per ray of the cell-Gauss volume adjoint: one per 55 voxels of the
// grid diagonal, within [1, 64] -- 8 on a 256^3 grid, measured best against
// 12 / 16 / 24 / 32 on a 360-view 512^2 full scan (joint VJP 418 / 423 / 429
// / 452 / 476 ms): each chunk is halved, up to three times, only while its
// brick overflows LDS, and fewer chunks mean fewer bounds and drains.
inline int gauss_chunks(int nx, int ny, int nz) {
    const double diagonal = std::sqrt((double)nx * nx + (double)ny * ny + (double)nz * nz);
    return (int)std::lround(std::fmin(64.0, std::fmax(1.0, diagonal / 55.0)));
}

// Samples per lockstep chunk of the volume adjoint: 32 at 1024 samples over
// a 256^3 grid (measured best against 16/24/48/64), scaled so that a chunk
// spans about the same distance in voxels -- 32 * (S / 1024) * (443 / d),
// d = the grid diagonal in voxels -- within [8, 64]: sparser sampling covers
// more voxels per sample and would otherwise grow the chunk bricks past LDS.
inline int chunk_samples(int samples, int nx, int ny, int nz) {
    const double diagonal = std::sqrt((double)nx * nx + (double)ny * ny + (double)nz * nz);
    const double chunk = 32.0 * samples / 1024.0 * 443.4 / diagonal;
    return (int)std::lround(std::fmin(64.0, std::fmax(8.0, chunk)));
}
// FP64 LDS brick entries per workgroup (dynamic LDS, 8 B each).  gfx950
// executes ds_add_f64 about ten times faster than ds_add_f32 (20 vs 193
// cycles per wave64 instruction, measured), so the brick accumulates the FP32
// cell sums in FP64.
constexpr int kBrickCapDefault = 4096;
// Shortest chunk the FP32 volume adjoint halves an overflowing chunk to
// before it flushes to device atomics.
constexpr int kMinChunk = 4;

// Host estimate of the chunk bricks at the identity pose, from up to four
// views spread over the scan (views 0, V/4, V/2, 3V/4): the grid samples
// touched by 16 x 16 pixel tiles (every 4th tile along each detector axis,
// at most 16 per axis) over one chunk of samples at a quarter, half and three
// quarters of the ray.  The counts of all sampled views are pooled.  A plan takes bricks of
// 2 * kBrickCapDefault FP64 entries (64 KiB of LDS, fewer workgroups per CU)
// only when at least a quarter of the sampled bricks overflow
// kBrickCapDefault yet fit the larger brick: magnified detectors over fine
// grids, whose tiles cover wide footprints and would otherwise fall back to
// device atomics.  (A 256^3 grid under a 512^2 panel at 1.5x magnification
// overflows in 9 % of the samples and keeps the default; a 512 x 512 x 139
// chest grid at 0.76 mm under a 512^2 panel at 2.67 mm overflows in 60 %.)
// The same samples decide whether the FP64 volume adjoint chunks by slabs
// (*slab_chunks): the chest grid's tiles at the silhouette overflow even the
// larger brick in 15 % of the index chunks and in none of the slab chunks.
int choose_brick_cap(const Params &P, const double *panels, int *slab_chunks) {
    const int n[3] = {P.nx, P.ny, P.nz};
    const int chunks = (P.samples + P.chunk_samples - 1) / P.chunk_samples;
    const int tile = 16;
    const int tiles_y = (P.rows + tile - 1) / tile, tiles_x = (P.cols + tile - 1) / tile;
    const int step_y = std::max(4, (tiles_y + 15) / 16), step_x = std::max(4, (tiles_x + 15) / 16);
    const int sampled_chunks[3] = {chunks / 4, chunks / 2, (3 * chunks) / 4};
    int measured = 0, rescued = 0, index_over = 0, slab_measured = 0, slab_over = 0;  // over: > 2 * default
    const int sampled_views = std::min(P.views, 4);
    for (int sv = 0; sv < sampled_views; ++sv)
    for (int ty = 0; ty < tiles_y; ty += step_y)
        for (int tx = 0; tx < tiles_x; tx += step_x) {
            const double *pv = panels + 12 * (size_t)((long long)sv * P.views / sampled_views);
            // the tile's rays in grid coordinates with their clipped intervals
            struct TileRay {
                double o[3], d[3], t0, step;
            };
            std::vector<TileRay> rays;
            for (int r = ty * tile; r < std::min(ty * tile + tile, P.rows); ++r)
                for (int c = tx * tile; c < std::min(tx * tile + tile, P.cols); ++c) {
                    TileRay tr;
                    double src[3], end[3];
                    for (int a = 0; a < 3; ++a) {
                        src[a] = pv[a] - P.go[a];
                        end[a] = pv[3 + a] + c * pv[6 + a] + r * pv[9 + a] - pv[a];
                    }
                    for (int a = 0; a < 3; ++a) {
                        tr.o[a] = P.g[3 * a] * src[0] + P.g[3 * a + 1] * src[1] + P.g[3 * a + 2] * src[2];
                        tr.d[a] = P.g[3 * a] * end[0] + P.g[3 * a + 1] * end[1] + P.g[3 * a + 2] * end[2];
                    }
                    double t0 = 0.0, t1 = 1.0;
                    for (int a = 0; a < 3; ++a) {
                        if (tr.d[a] == 0.0) {
                            if (tr.o[a] < -0.5 || tr.o[a] > n[a] - 0.5) t1 = -1.0;
                            continue;
                        }
                        const double u = (-0.5 - tr.o[a]) / tr.d[a], v = (n[a] - 0.5 - tr.o[a]) / tr.d[a];
                        t0 = std::max(t0, std::min(u, v));
                        t1 = std::min(t1, std::max(u, v));
                    }
                    if (!(t1 > t0)) continue;
                    tr.t0 = t0;
                    tr.step = (t1 - t0) / P.samples;
                    rays.push_back(tr);
                }
            if (rays.empty()) continue;
            // slab chunks (the device's slab_setup / slab_range, in FP64 on the host)
            int votes[3] = {0, 0, 0};
            for (const TileRay &tr : rays) {
                const double ad[3] = {std::fabs(tr.d[0]), std::fabs(tr.d[1]), std::fabs(tr.d[2])};
                ++votes[ad[0] >= ad[1] && ad[0] >= ad[2] ? 0 : (ad[1] >= ad[2] ? 1 : 2)];
            }
            const int A = votes[0] >= votes[1] && votes[0] >= votes[2] ? 0 : (votes[1] >= votes[2] ? 1 : 2);
            int slo = INT_MAX, shi = INT_MIN, adv = 1;
            for (const TileRay &tr : rays) {
                const double q0 = tr.o[A] + (tr.t0 + 0.5 * tr.step) * tr.d[A];
                const double q1 = tr.o[A] + (tr.t0 + (P.samples - 0.5) * tr.step) * tr.d[A];
                slo = std::min(slo, (int)std::floor(std::min(std::max(std::min(q0, q1), -2.0), n[A] + 1.0)));
                shi = std::max(shi, (int)std::ceil(std::min(std::max(std::max(q0, q1), -2.0), n[A] + 1.0)));
                adv = std::max(adv, (int)std::min(std::ceil(tr.step * std::fabs(tr.d[A]) * 64.0), 1048576.0));
            }
            const int width = std::max(1, (P.chunk_samples * adv + 63) / 64);
            const int count = std::max(1, (shi - slo + width - 1) / width);
            for (int j : {count / 4, count / 2, (3 * count) / 4}) {
                int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
                for (const TileRay &tr : rays) {
                    const double da = tr.d[A];
                    if (da == 0.0) continue;
                    const int js = da > 0.0 ? j : count - 1 - j;
                    const double b0 = slo + js * width, b1 = slo + (js + 1) * width;
                    const double t0 = ((b0 - tr.o[A]) / da - tr.t0) / tr.step - 0.5;
                    const double t1 = ((b1 - tr.o[A]) / da - tr.t0) / tr.step - 0.5;
                    double k0, k1;
                    if (da > 0.0) {
                        k0 = j == 0 ? 0.0 : std::min(std::max(std::ceil(t0), 0.0), (double)P.samples);
                        k1 = j == count - 1 ? P.samples : std::min(std::max(std::ceil(t1), 0.0), (double)P.samples);
                    } else {
                        k0 = j == 0 ? 0.0 : std::min(std::max(std::floor(t1) + 1.0, 0.0), (double)P.samples);
                        k1 = j == count - 1 ? P.samples : std::min(std::max(std::floor(t0) + 1.0, 0.0), (double)P.samples);
                    }
                    if (!(k1 > k0)) continue;
                    for (double k : {k0, k1 - 1.0}) {
                        const double pos = tr.t0 + (k + 0.5) * tr.step;
                        for (int a = 0; a < 3; ++a) {
                            const double x = std::min(std::max(tr.o[a] + pos * tr.d[a], 0.0), (double)(n[a] - 1));
                            const int l = std::min((int)std::floor(x), std::max(n[a] - 2, 0));
                            lo[a] = std::min(lo[a], l);
                            hi[a] = std::max(hi[a], l);
                        }
                    }
                }
                if (lo[0] == INT_MAX) continue;
                long long volume = 1;
                for (int a = 0; a < 3; ++a) volume *= std::min(hi[a] + 1, n[a] - 1) - lo[a] + 1;
                ++slab_measured;
                slab_over += volume > 2 * kBrickCapDefault;
            }
            for (int q = 0; q < 3; ++q) {
                const int k0 = sampled_chunks[q] * P.chunk_samples, k1 = std::min(k0 + P.chunk_samples, P.samples);
                int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
                for (int r = ty * tile; r < std::min(ty * tile + tile, P.rows); ++r)
                    for (int c = tx * tile; c < std::min(tx * tile + tile, P.cols); ++c) {
                        double src[3], end[3], o[3], d[3];
                        for (int a = 0; a < 3; ++a) {
                            src[a] = pv[a] - P.go[a];
                            end[a] = pv[3 + a] + c * pv[6 + a] + r * pv[9 + a] - pv[a];
                        }
                        for (int a = 0; a < 3; ++a) {
                            o[a] = P.g[3 * a] * src[0] + P.g[3 * a + 1] * src[1] + P.g[3 * a + 2] * src[2];
                            d[a] = P.g[3 * a] * end[0] + P.g[3 * a + 1] * end[1] + P.g[3 * a + 2] * end[2];
                        }
                        double t0 = 0.0, t1 = 1.0;
                        for (int a = 0; a < 3; ++a) {
                            if (d[a] == 0.0) {
                                if (o[a] < -0.5 || o[a] > n[a] - 0.5) t1 = -1.0;
                                continue;
                            }
                            const double u = (-0.5 - o[a]) / d[a], v = (n[a] - 0.5 - o[a]) / d[a];
                            t0 = std::max(t0, std::min(u, v));
                            t1 = std::min(t1, std::max(u, v));
                        }
                        if (!(t1 > t0)) continue;
                        const double step = (t1 - t0) / P.samples;
                        for (int k : {k0, k1 - 1}) {
                            const double pos = t0 + (k + 0.5) * step;
                            for (int a = 0; a < 3; ++a) {
                                const double x = std::min(std::max(o[a] + pos * d[a], 0.0), (double)(n[a] - 1));
                                const int l = std::min((int)std::floor(x), std::max(n[a] - 2, 0));
                                lo[a] = std::min(lo[a], l);
                                hi[a] = std::max(hi[a], l);
                            }
                        }
                    }
                if (lo[0] == INT_MAX) continue;
                long long volume = 1;
                for (int a = 0; a < 3; ++a) volume *= std::min(hi[a] + 1, n[a] - 1) - lo[a] + 1;
                ++measured;
                rescued += volume > kBrickCapDefault && volume <= 2 * kBrickCapDefault;
                index_over += volume > 2 * kBrickCapDefault;
            }
        }
    // Slab chunks cost lockstep iterations (a wave walks each slab as long as
    // its lane with the most samples there), so a plan takes them only when
    // index chunks overflow even the larger brick in at least 1/16 of the
    // samples and slab chunks in less than half as many: then with the
    // larger brick.
    *slab_chunks = measured > 0 && slab_measured > 0 && 16 * index_over >= measured &&
                   2 * slab_over * measured < index_over * slab_measured;
    if (*slab_chunks) return 2 * kBrickCapDefault;
    return measured > 0 && 4 * rescued >= measured ? 2 * kBrickCapDefault : kBrickCapDefault;
}

// Clamped lower cell coordinate of one FP32 sample (same arithmetic as cell32,
// without the support test: out-of-support samples only widen the box).
__device__ __forceinline__ void lower_cell32(const Params &P, float p0, float p1, float p2, int *l) {
    l[0] = min((int)floorf(fminf(fmaxf(p0, 0.0f), P.cx)), P.capx);
    l[1] = min((int)floorf(fminf(fmaxf(p1, 0.0f), P.cy)), P.capy);
    l[2] = min((int)floorf(fminf(fmaxf(p2, 0.0f), P.cz)), P.capz);
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

// Bound one chunk's cells over the workgroup and, when the box fits the LDS
// brick, describe it.  lo / nhi: this lane's lower cell corner minimum and
// negated maximum over its chunk samples (INT_MAX when it has none).  Every
// lane gets the same answer (kept in SGPRs); one barrier.
__device__ __forceinline__ bool chunk_brick(const Params &P, int (*red)[6], int *lo, int *nhi, bool unit_x,
                                            Brick &B, bool *overflow = nullptr, bool pad = false) {
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        nhi[a] = wave_min(nhi[a]);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            red[wv][a] = lo[a];
            red[wv][3 + a] = nhi[a];
        }
    }
    __syncthreads();
    int l[3], h[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        l[a] = min(min(red[0][a], red[1][a]), min(red[2][a], red[3][a]));
        h[a] = -min(min(red[0][3 + a], red[1][3 + a]), min(red[2][3 + a], red[3][3 + a]));
    }
    if (l[0] == INT_MAX) return false;
    const int n[3] = {P.nx, P.ny, P.nz};
    int e[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) e[a] = min(h[a] + 1, n[a] - 1) - l[a] + 1;
    const long long vol = (long long)e[0] * e[1] * e[2];
    if (overflow != nullptr) *overflow = vol > P.brick_cap;
    if (vol > P.brick_cap) return false;
    // every lane holds the same values: keep them in SGPRs
    B.x0 = __builtin_amdgcn_readfirstlane(l[0]);
    B.y0 = __builtin_amdgcn_readfirstlane(l[1]);
    B.z0 = __builtin_amdgcn_readfirstlane(l[2]);
    // pad: odd row and plane strides when the padded box still fits (FP64
    // entries whose indices differ by a multiple of 32 share an LDS bank, so
    // even strides line up the corners of neighbouring rays' cells).  Only
    // for callers that write cells inside the bound box: the padding stays
    // zero and the drain skips it.
    int px = e[0], py = e[1];
    // Padded strides keep flush_lds's global-coordinate index below 2^24 per
    // operand and 2^31 overall when nz (ny + 1) and nz (ny + 1) (nx + 1) do.
    const long long plane = (long long)P.nz * (P.ny + 1);
    if (pad && (long long)(px | 1) * (py | 1) * e[2] <= P.brick_cap && plane <= (1 << 24) &&
        plane * (P.nx + 1) <= INT_MAX)
        px |= 1, py |= 1;
    B.bx = __builtin_amdgcn_readfirstlane(px);
    B.by = __builtin_amdgcn_readfirstlane(py);
    B.bz = __builtin_amdgcn_readfirstlane(e[2]);
    B.lsx = (unit_x || P.nx > 1) ? 1 : 0;
    B.lsy = P.ny > 1 ? B.bx : 0;
    B.lsz = P.nz > 1 ? B.bx * B.by : 0;
    B.volume = B.bx * B.by * B.bz;
    B.origin = (B.z0 * B.by + B.y0) * B.bx + B.x0;
    return true;
}

// Add the brick's non-zero entries to the FP64 accumulator (one device
// atomic per touched grid sample) and clear them.  Call after a barrier.
__device__ __forceinline__ void drain_brick(const Params &P, const Brick &B, double *brick, double *acc) {
    const float inv_bx = 1.0f / (float)B.bx, inv_by = 1.0f / (float)B.by;
    for (int i = threadIdx.x; i < B.volume; i += kBlock) {
        const double v = brick[i];
        if (v != 0.0) {
            brick[i] = 0.0;
            // i = (z * by + y) * bx + x: float quotients corrected by one step
            int q = (int)((float)i * inv_bx);
            q -= (q * B.bx > i) ? 1 : 0;
            q += ((q + 1) * B.bx <= i) ? 1 : 0;
            const int x = i - q * B.bx;
            int z = (int)((float)q * inv_by);
            z -= (z * B.by > q) ? 1 : 0;
            z += ((z + 1) * B.by <= q) ? 1 : 0;
            const int y = q - z * B.by;
            const int g = ((B.z0 + z) * P.ny + (B.y0 + y)) * P.nx + (B.x0 + x);
            unsafeAtomicAdd(&acc[g], v);
        }
    }
}

// ---------------------------------------------------------- FP32 VJP ------
// kernels/projection_fp32.py:377-617.  The workgroup walks its rays' samples
// in lockstep chunks.  Per chunk it bounds the cells its rays touch (sample
// positions are monotone in the sample index, so the first and last sample of
// the chunk bound each ray), privatises that brick in LDS when it fits, sums
// register cell accumulators into it with LDS atomics, then adds the non-zero
// brick entries to the FP64 accumulator with one global atomic each.  Chunks
// whose brick does not fit flush cells with global atomics directly.
//
// Volume factors f = seed * length * step.  Rays with 2^-102 <= |f| <= 2^96
// accumulate RN(fma(w, f, s)) per cell in the sample loop: against the
// reference's RN(f * w) summed in FP32 this drops one rounding per sample,
// and the reference's flush of increments below 2^-126
// (kernels/projection_fp32.py:240-252) can only concern terms with
// w < 2^-24, each below 2^-126; the bound 2^96 keeps every FP32 cell sum
// finite.  Other rays (smaller factors, whose increments the reference
// flushes to exact zeros, and larger or non-finite ones) add the reference's
// per-increment RN(f * w) (flushed, or unflushed for an FP64 destination)
// with device atomics after the pose epilogue.
//
// The sample loop is written for wave-level execution: some lane of a wave
// changes cell at nearly every sample, so the cell-change branch (corner
// loads, one cell flush, register reset) runs once per sample for the wave,
// and everything the loop carries is updated in place.  Samples outside the
// support (only waves whose chunk ends leave it test per sample) take the
// clamped cell with zero fractions, a zero factor and zero value and slopes:
// they add nothing, as the reference skips them.


template <typename T, bool VOL, bool POSE, int C, bool COMPENSATED, int LAYOUT, bool PIPE = false>
__global__ __launch_bounds__(kBlock) void vjp32_kernel(Params P, const T *__restrict__ field,
                                                       const float *__restrict__ seeds,
                                                       double *__restrict__ acc,
                                                       double *__restrict__ partials) {
    extern __shared__ double brick[];
    __shared__ int bound_red[kBlock / 64][6];
    int view, row, col, wslot;
    bool wave_valid;
    if (!lane_ray(P, view, row, col, wslot, wave_valid)) return;
    const double *pose = P.poses + 12 * (size_t)view;
    // this workgroup's samples [first, last) and its pose partial slot
    int first = 0, last = P.samples;
    if (P.segments > 1) {
        const int b = blockIdx.x;
        const int seg = ((b % kXcds) * P.chunk + b / kXcds) / P.base_tiles;
        first = seg * P.segment_samples;
        last = min(first + P.segment_samples, P.samples);
        wslot += (view * (P.segments - 1) + seg) * P.wt_per_view;
    }

    // ---- per-ray setup
    bool active = false, summed = false, scattered = false, bad = false;
    float o[3] = {0, 0, 0}, d[3] = {0, 0, 0}, len = 0.0f, seed = 0.0f, span = 0.0f, step = 0.0f;
    float volume_factor = 0.0f;
    Interval32 iv{};
    if (ray_in_range(P, view, row, col)) {
        const size_t ray = ((size_t)view * P.rows + row) * P.cols + col;
        World w;
        ray_world(P, pose, view, row, col, w);
        Ray r;
        ray_grid(P, w, r);
        o[0] = (float)r.o[0];
        o[1] = (float)r.o[1];
        o[2] = (float)r.o[2];
        d[0] = (float)r.d[0];
        d[1] = (float)r.d[1];
        d[2] = (float)r.d[2];
        len = (float)r.len;
        if (!r.valid || !finite3f(o[0], o[1], o[2]) || !finite3f(d[0], d[1], d[2]) || !isfinite(len) ||
            len <= 0.0f) {
            flag(P, view, DPT_STATUS_RANGE);
        } else {
            iv = interval32(P, o, d);
            span = iv.upper - iv.lower;
            seed = seeds[ray];
            if (!isfinite(seed)) flag(P, view, DPT_STATUS_RANGE);
            if (span > 0.0f && seed != 0.0f) {
                active = true;
                step = span / (float)P.samples;
                volume_factor = seed * (len * step);
                // sample points are monotone in the sample index: all are
                // finite iff the first and the last are
                const bool finite = finite_span(o, d, iv.lower, step, 0, P.samples);
                bad = !finite;
                const float m = fabsf(volume_factor);
                summed = VOL && finite && m >= 0x1p-102f && m <= 0x1p96f;  // false for NaN
                scattered = VOL && !summed;
            }
        }
    }
    // the factor of the in-loop cell sums (0 for rays that add elsewhere)
    const float loop_factor = summed ? volume_factor : 0.0f;

    Acc<COMPENSATED> mean, dor0, dor1, dor2, dd0, dd1, dd2, shl, shu;
    mean.init();
    dor0.init();
    dor1.init();
    dor2.init();
    dd0.init();
    dd1.init();
    dd2.init();
    shl.init();
    shu.init();
    // cs: the current cell's sums.  Volume-only launches keep a departed
    // cell's sums in ps until some lane of the wave leaves a cell while
    // still holding one, so that a wave's flushes gather more lanes
    // (volume-only VJP 190.7 -> 175.5 ms at the full scan; the joint kernel,
    // short of registers, ran slower with it and flushes at once).
    constexpr bool DEFER = VOL && !POSE;
    CellSums cs, ps;
    cs.reset();
    ps.reset();

    static_assert(!PIPE || (POSE && !VOL && !COMPENSATED), "the pipeline is a pose-only FP32 loop");
    if constexpr (VOL) {
        for (int i = threadIdx.x; i < P.brick_cap; i += kBlock) brick[i] = 0.0;
    }
    if constexpr (PIPE) {
        if (active) {
            // slot j holds sample k's corner values, fractions, slopes,
            // position and support flag from its issue kPipeDepth samples back
            float ring[kPipeDepth][8], rf[kPipeDepth][6], rpos[kPipeDepth], rlive[kPipeDepth];
            auto issue = [&](int j, int k) {
                const float pos = fmaf(sample_centre(P, k), step, iv.lower);
                const float p0 = fmaf(pos, d[0], o[0]);
                const float p1 = fmaf(pos, d[1], o[1]);
                const float p2 = fmaf(pos, d[2], o[2]);
                Cell32 c;
                cell32_inside(P, p0, p1, p2, c);  // clamped: the load stays in the field
                const bool inside = inside32(P, p0, p1, p2);
                rf[j][0] = inside ? c.fx : 0.0f;
                rf[j][1] = inside ? c.fy : 0.0f;
                rf[j][2] = inside ? c.fz : 0.0f;
                rf[j][3] = inside ? c.slx : 0.0f;
                rf[j][4] = inside ? c.sly : 0.0f;
                rf[j][5] = inside ? c.slz : 0.0f;
                rpos[j] = pos;
                rlive[j] = inside ? 1.0f : 0.0f;
                load8<T, LAYOUT>(P, field, c.base, ring[j]);
            };
            auto consume = [&](int j, int k) {
                Cell32 c;
                c.fx = rf[j][0];
                c.fy = rf[j][1];
                c.fz = rf[j][2];
                c.slx = rf[j][3];
                c.sly = rf[j][4];
                c.slz = rf[j][5];
                float wt[8], g[3];
                weights32(c, wt);
                const float value = interpolate32(wt, ring[j]) * rlive[j];
                slopes32(c, ring[j], g);
                const float pos = rpos[j];
                mean.add(value);
                dor0.add(g[0]);
                dor1.add(g[1]);
                dor2.add(g[2]);
                dd0.fma_add(pos, g[0]);
                dd1.fma_add(pos, g[1]);
                dd2.fma_add(pos, g[2]);
                const float alpha = sample_alpha(P, k);
                const float along = fmaf(g[2], d[2], fmaf(g[1], d[1], g[0] * d[0]));
                shl.fma_add(1.0f - alpha, along);
                shu.fma_add(alpha, along);
            };
#pragma unroll
            for (int j = 0; j < kPipeDepth; ++j) issue(j, min(first + j, last - 1));
            // whole groups without branches, so that the loads stay in flight
            // across iterations; the last group's issues repeat sample
            // last - 1 and are not consumed
            int k0 = first;
            for (; k0 + kPipeDepth <= last; k0 += kPipeDepth) {
#pragma unroll
                for (int j = 0; j < kPipeDepth; ++j) {
                    consume(j, k0 + j);
                    issue(j, min(k0 + j + kPipeDepth, last - 1));
                }
            }
#pragma unroll
            for (int j = 0; j < kPipeDepth; ++j)
                if (k0 + j < last) consume(j, k0 + j);  // uniform
        }
    } else {
    // One sample of the walk: the cell, its corner values on a cell change
    // (with the previous cell's flush), the volume cell sums and the pose
    // sums.  SLAB: per-lane sample indices (table reads by vector loads) and
    // support tests on every sample.
    auto take = [&](auto slab, int k, bool checked, const Brick &B, bool use_lds, int &cur, float *v) {
        constexpr bool SLAB = decltype(slab)::value;
        const float pos = fmaf(SLAB ? (float)k + 0.5f : sample_centre(P, k), step, iv.lower);
        const float p0 = fmaf(pos, d[0], o[0]);
        const float p1 = fmaf(pos, d[1], o[1]);
        const float p2 = fmaf(pos, d[2], o[2]);
        Cell32 c;
        cell32_inside(P, p0, p1, p2, c);
        float factor = loop_factor, live = 1.0f;
        if (SLAB || checked) {  // outside samples add nothing
            const bool inside = inside32(P, p0, p1, p2);
            c.fx = inside ? c.fx : 0.0f;
            c.fy = inside ? c.fy : 0.0f;
            c.fz = inside ? c.fz : 0.0f;
            c.slx = inside ? c.slx : 0.0f;
            c.sly = inside ? c.sly : 0.0f;
            c.slz = inside ? c.slz : 0.0f;
            factor = inside ? factor : 0.0f;
            live = inside ? 1.0f : 0.0f;
        }
        const bool changed = c.base != cur;
        if constexpr (DEFER) {
            if (__builtin_amdgcn_ballot_w64(changed && ps.x >= 0) != 0) {  // uniform
                if (ps.x >= 0) {
                    if (use_lds)
                        flush_lds(B, ps, brick);
                    else
                        flush_device(P, ps, acc);
                    ps.x = -1;
                }
            }
        }
        if (changed) {
            if constexpr (POSE) load8<T, LAYOUT>(P, field, c.base, v);
            if constexpr (DEFER) {
                ps = cs;
            } else if constexpr (VOL) {
                if (cs.x >= 0) {
                    if (use_lds)
                        flush_lds(B, cs, brick);
                    else
                        flush_device(P, cs, acc);
                }
            }
            if constexpr (VOL) {
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) cs.s[kk] = 0.0f;
                // lanes that add elsewhere keep an empty cell
                cs.x = summed ? c.lx : -1;
                cs.y = c.ly;
                cs.z = c.lz;
            }
            cur = c.base;
        }
        float wt[8];
        weights32(c, wt);
        if constexpr (VOL) {
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) cs.s[kk] = fmaf(wt[kk], factor, cs.s[kk]);
        }
        if constexpr (POSE) {
            float g[3];
            // an outside sample's slopes are zero (above); its value
            // is zeroed here
            const float value = interpolate32(wt, v) * live;
            slopes32(c, v, g);
            mean.add(value);
            dor0.add(g[0]);
            dor1.add(g[1]);
            dor2.add(g[2]);
            // contracted by the reference's code generator (fuse_fp)
            dd0.fma_add(pos, g[0]);
            dd1.fma_add(pos, g[1]);
            dd2.fma_add(pos, g[2]);
            const float alpha = SLAB ? P.alpha[k] : sample_alpha(P, k);
            const float along = fmaf(g[2], d[2], fmaf(g[1], d[1], g[0] * d[0]));
            shl.fma_add(1.0f - alpha, along);
            shu.fma_add(alpha, along);
        }
    };
    auto fold_all = [&]() {
        if constexpr (POSE && COMPENSATED) {
            mean.fold();
            dor0.fold();
            dor1.fold();
            dor2.fold();
            dd0.fold();
            dd1.fold();
            dd2.fold();
            shl.fold();
            shu.fold();
        }
    };
    // Chunks of P.chunk_samples samples; a chunk whose brick overflows LDS is
    // halved (down to kMinChunk samples) until it fits; if it still overflows
    // the workgroup walks its remaining samples by slab chunks (below), or,
    // without the fallback, flushes that chunk to device atomics.  Chunk
    // bounds are uniform over the workgroup.
    int switch_at = -1;
    for (int k0 = first, k1 = first; k0 < last; k0 = k1) {
        Brick B{};
        bool use_lds = false;
        k1 = min(k0 + P.chunk_samples, last);
        if constexpr (VOL) {
            bool overflow = false;
            for (int len = P.chunk_samples;; len >>= 1) {
                k1 = min(k0 + len, last);
                // ---- bound the cells of [k0, k1) over the workgroup
                int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, nhi[3] = {INT_MAX, INT_MAX, INT_MAX};
                if (active && summed) {
                    const float pa = fmaf((float)k0 + 0.5f, step, iv.lower);
                    const float pb = fmaf((float)(k1 - 1) + 0.5f, step, iv.lower);
                    int la[3], lb[3];
                    lower_cell32(P, fmaf(pa, d[0], o[0]), fmaf(pa, d[1], o[1]), fmaf(pa, d[2], o[2]), la);
                    lower_cell32(P, fmaf(pb, d[0], o[0]), fmaf(pb, d[1], o[1]), fmaf(pb, d[2], o[2]), lb);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = min(la[a], lb[a]);
                        nhi[a] = -max(la[a], lb[a]);
                    }
                }
                // x-pair and cell layouts run only on grids with nx > 1: a
                // compile-time unit x stride lets the LDS atomics of
                // x-adjacent corners use the instruction's offset field
                use_lds = chunk_brick(P, bound_red, lo, nhi, LAYOUT != kLayoutScalar, B, &overflow, true);
                if (!overflow || len <= kMinChunk) break;
                __syncthreads();  // the next bound rewrites bound_red
            }
            if (overflow && P.slab_fallback) {  // uniform
                switch_at = k0;
                break;
            }
        }
        if (active) {
            float v[8];
            int cur = -1;
            // per-sample support tests only for waves with a ray whose chunk
            // ends leave the support (uniform over the wave)
            const bool checked = __builtin_amdgcn_ballot_w64(!span_inside(P, o, d, iv.lower, step, k0, k1)) != 0;
            for (int k = k0; k < k1; ++k) take(std::false_type{}, k, checked, B, use_lds, cur, v);
        }
        fold_all();
        if constexpr (VOL) {
            if (DEFER && ps.x >= 0) {
                if (use_lds)
                    flush_lds(B, ps, brick);
                else
                    flush_device(P, ps, acc);
            }
            if (cs.x >= 0) {
                if (use_lds)
                    flush_lds(B, cs, brick);
                else
                    flush_device(P, cs, acc);
            }
            ps.reset();
            cs.reset();
            if (use_lds) {
                __syncthreads();
                drain_brick(P, B, brick, acc);
            }
            __syncthreads();
        }
    }
    if constexpr (VOL) {
        if (switch_at >= 0) {  // uniform
            // ---- slab chunks over [switch_at, last): space cut into slabs of
            // `width` grid samples along the workgroup's dominant axis; each
            // lane walks, per slab, its samples whose axis coordinate lies in
            // the slab.  Boundaries are integers and both sides of one use
            // the same expression, so every sample falls in exactly one slab,
            // and a lane's slabs come in its sample order (its pose sums keep
            // their order).  The brick bound comes from each slab's own end
            // samples.
            __syncthreads();  // the last chunk bound is read
            const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
            const float ad[3] = {fabsf(d[0]), fabsf(d[1]), fabsf(d[2])};
            const int mine = ad[0] >= ad[1] && ad[0] >= ad[2] ? 0 : (ad[1] >= ad[2] ? 1 : 2);
            int my_votes[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) my_votes[a] = __popcll(__builtin_amdgcn_ballot_w64(active && mine == a));
            if (lane == 0)
                for (int a = 0; a < 3; ++a) bound_red[wv][a] = my_votes[a];
            __syncthreads();
            int votes[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
                votes[a] = bound_red[0][a] + bound_red[1][a] + bound_red[2][a] + bound_red[3][a];
            const int axis = votes[0] >= votes[1] && votes[0] >= votes[2] ? 0 : (votes[1] >= votes[2] ? 1 : 2);
            const int n_axis = axis == 0 ? P.nx : (axis == 1 ? P.ny : P.nz);
            const float oa = axis == 0 ? o[0] : (axis == 1 ? o[1] : o[2]);
            const float da = axis == 0 ? d[0] : (axis == 1 ? d[1] : d[2]);
            int slo = INT_MAX, snhi = INT_MAX, nadv = INT_MAX;
            if (active) {
                const float q0 = fmaf(fmaf((float)switch_at + 0.5f, step, iv.lower), da, oa);
                const float q1 = fmaf(fmaf((float)(last - 1) + 0.5f, step, iv.lower), da, oa);
                slo = (int)floorf(fminf(fmaxf(fminf(q0, q1), -2.0f), (float)n_axis + 1.0f));
                snhi = -(int)ceilf(fminf(fmaxf(fmaxf(q0, q1), -2.0f), (float)n_axis + 1.0f));
                nadv = -(int)fminf(fmaxf(ceilf(step * fabsf(da) * 64.0f), 1.0f), 1048576.0f);
            }
            slo = wave_min(slo);
            snhi = wave_min(snhi);
            nadv = wave_min(nadv);
            __syncthreads();  // every wave has read the votes
            if (lane == 0) {
                bound_red[wv][3] = slo;
                bound_red[wv][4] = snhi;
                bound_red[wv][5] = nadv;
            }
            __syncthreads();
            int slab_lo = min(min(bound_red[0][3], bound_red[1][3]), min(bound_red[2][3], bound_red[3][3]));
            const int slab_hi = -min(min(bound_red[0][4], bound_red[1][4]), min(bound_red[2][4], bound_red[3][4]));
            const int adv = -min(min(bound_red[0][5], bound_red[1][5]), min(bound_red[2][5], bound_red[3][5]));
            __syncthreads();  // bound_red is reused by the slab bounds
            int width = 1, count = 1;
            if (slab_lo == INT_MAX) {
                slab_lo = 0;
            } else {
                width = max(1, (P.chunk_samples * adv + 63) / 64);
                count = max(1, (slab_hi - slab_lo + width - 1) / width);
            }
            for (int ch = 0; ch < count; ++ch) {
                int k0 = switch_at, k1 = switch_at;
                if (active) {
                    if (da == 0.0f) {
                        const int own = min(max((int)floorf((oa - (float)slab_lo) / (float)width), 0), count - 1);
                        k1 = ch == own ? last : switch_at;
                    } else {
                        // t(b): the sample index where the coordinate crosses b
                        const int js = da > 0.0f ? ch : count - 1 - ch;
                        const float b0 = (float)(slab_lo + js * width), b1 = (float)(slab_lo + (js + 1) * width);
                        const float t0 = ((b0 - oa) / da - iv.lower) / step - 0.5f;
                        const float t1 = ((b1 - oa) / da - iv.lower) / step - 0.5f;
                        const float lo_k = (float)switch_at, hi_k = (float)last;
                        if (da > 0.0f) {
                            k0 = ch == 0 ? switch_at : (int)fminf(fmaxf(ceilf(t0), lo_k), hi_k);
                            k1 = ch == count - 1 ? last : (int)fminf(fmaxf(ceilf(t1), lo_k), hi_k);
                        } else {
                            k0 = ch == 0 ? switch_at : (int)fminf(fmaxf(floorf(t1) + 1.0f, lo_k), hi_k);
                            k1 = ch == count - 1 ? last : (int)fminf(fmaxf(floorf(t0) + 1.0f, lo_k), hi_k);
                        }
                        k1 = max(k1, k0);
                    }
                }
                int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, nhi[3] = {INT_MAX, INT_MAX, INT_MAX};
                if (summed && k1 > k0) {
                    const float pa = fmaf((float)k0 + 0.5f, step, iv.lower);
                    const float pb = fmaf((float)(k1 - 1) + 0.5f, step, iv.lower);
                    int la[3], lb[3];
                    lower_cell32(P, fmaf(pa, d[0], o[0]), fmaf(pa, d[1], o[1]), fmaf(pa, d[2], o[2]), la);
                    lower_cell32(P, fmaf(pb, d[0], o[0]), fmaf(pb, d[1], o[1]), fmaf(pb, d[2], o[2]), lb);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = min(la[a], lb[a]);
                        nhi[a] = -max(la[a], lb[a]);
                    }
                }
                Brick B{};
                const bool use_lds = chunk_brick(P, bound_red, lo, nhi, LAYOUT != kLayoutScalar, B, nullptr, true);
                if (active) {
                    float v[8];
                    int cur = -1;
                    for (int k = k0; k < k1; ++k) take(std::true_type{}, k, true, B, use_lds, cur, v);
                }
                fold_all();
                if (DEFER && ps.x >= 0) {
                    if (use_lds)
                        flush_lds(B, ps, brick);
                    else
                        flush_device(P, ps, acc);
                }
                if (cs.x >= 0) {
                    if (use_lds)
                        flush_lds(B, cs, brick);
                    else
                        flush_device(P, cs, acc);
                }
                ps.reset();
                cs.reset();
                __syncthreads();
                if (use_lds) drain_brick(P, B, brick, acc);
                __syncthreads();
            }
        }
    }
    }
    if constexpr (POSE) {
        double gradient[kMaxComponents];
#pragma unroll
        for (int c = 0; c < kMaxComponents; ++c) gradient[c] = 0.0;
        if (active) {
            const float m = mean.value();
            const float lower_seed = fmaf(span, shl.value(), -m);
            const float upper_seed = fmaf(span, shu.value(), m);
            const float dor[3] = {dor0.value(), dor1.value(), dor2.value()};
            const float ddir[3] = {dd0.value(), dd1.value(), dd2.value()};
            double og[3], dg[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float lo_o = (iv.lo_axis == a) ? iv.lo_o : 0.0f;
                const float lo_d = (iv.lo_axis == a) ? iv.lo_d : 0.0f;
                const float up_o = (iv.up_axis == a) ? iv.up_o : 0.0f;
                const float up_d = (iv.up_axis == a) ? iv.up_d : 0.0f;
                og[a] = (double)fmaf(upper_seed, up_o, fmaf(lower_seed, lo_o, span * dor[a]));
                dg[a] = (double)fmaf(upper_seed, up_d, fmaf(lower_seed, lo_d, span * ddir[a]));
            }
            const float factor = (seed * len) / (float)P.samples;
            World w;  // recomputed: keeps 24 FP64 registers out of the sample loop
            ray_world(P, pose, view, row, col, w);
            pose_epilogue<C>(P, pose, w, (double)factor, og, dg, gradient, (double)seed, (double)(m * step));
        }
        store_gradient<C>(P, gradient, partials, wslot, wave_valid, view, row, col);
    }
    if constexpr (VOL) {
        // Rays outside the in-loop factor range: the reference's per-increment
        // RN(f * w) (kernels/projection_fp32.py:240-252, 513-539) per sample
        // and corner, flushed below 2^-126 for an FP32 destination and kept
        // for an FP64 one (kernels/projection_fp32.py:212-235), with device
        // atomics (zeros skipped).
        if (scattered) {
            for (int k = first; k < last; ++k) {
                const float pos = fmaf(sample_centre(P, k), step, iv.lower);
                const float p0 = fmaf(pos, d[0], o[0]);
                const float p1 = fmaf(pos, d[1], o[1]);
                const float p2 = fmaf(pos, d[2], o[2]);
                if (!inside32(P, p0, p1, p2)) continue;
                Cell32 c;
                cell32_inside(P, p0, p1, p2, c);
                float wt[8];
                weights32(c, wt);
                const int voxel = (c.lz * P.ny + c.ly) * P.nx + c.lx;
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    const float inc = P.volume_f64 ? volume_factor * wt[kk] : scatter_increment(volume_factor, wt[kk]);
                    bad |= !isfinite(inc);
                    if (inc != 0.0f) unsafeAtomicAdd(&acc[voxel + corner_offset(P, kk)], (double)inc);
                }
            }
        }
    }
    if (bad) flag(P, view, DPT_STATUS_RANGE);
}

// FP32 volume adjoint by slab chunks (plans whose index chunks overflow the
// brick: tiles at the grid's silhouette, where the clipped intervals of a
// tile's rays differ in length and their k-th samples spread apart).  The
// workgroup cuts its samples [first, last) into slabs of W grid samples
// along its dominant axis A and each lane visits, per chunk, its samples
// whose A coordinate lies in the slab (W: the advance of the fastest lane
// over the plan's chunk length).  Boundaries are integers and both sides of
// one use the same expression, so every sample falls in exactly one chunk;
// the brick bound comes from the chunk's own end samples, so any rounding in
// the inversion only moves samples between neighbouring chunks.  Per-sample
// arithmetic, cell sums and flushes are those of vjp32_kernel's volume part.
__global__ __launch_bounds__(kBlock) void vol32_slab_kernel(Params P, const float *__restrict__ seeds,
                                                            double *__restrict__ acc) {
    extern __shared__ double brick[];
    __shared__ int bound_red[kBlock / 64][6];
    int view, row, col, wslot;
    bool wave_valid;
    if (!lane_ray(P, view, row, col, wslot, wave_valid)) return;
    const double *pose = P.poses + 12 * (size_t)view;
    int first = 0, last = P.samples;
    if (P.segments > 1) {
        const int b = blockIdx.x;
        const int seg = ((b % kXcds) * P.chunk + b / kXcds) / P.base_tiles;
        first = seg * P.segment_samples;
        last = min(first + P.segment_samples, P.samples);
    }
    bool active = false, summed = false, scattered = false, bad = false;
    float o[3] = {0, 0, 0}, d[3] = {0, 0, 0}, step = 0.0f, volume_factor = 0.0f;
    Interval32 iv{};
    if (ray_in_range(P, view, row, col)) {
        const size_t ray = ((size_t)view * P.rows + row) * P.cols + col;
        World w;
        ray_world(P, pose, view, row, col, w);
        Ray r;
        ray_grid(P, w, r);
        o[0] = (float)r.o[0];
        o[1] = (float)r.o[1];
        o[2] = (float)r.o[2];
        d[0] = (float)r.d[0];
        d[1] = (float)r.d[1];
        d[2] = (float)r.d[2];
        const float len = (float)r.len;
        if (!r.valid || !finite3f(o[0], o[1], o[2]) || !finite3f(d[0], d[1], d[2]) || !isfinite(len) ||
            len <= 0.0f) {
            flag(P, view, DPT_STATUS_RANGE);
        } else {
            iv = interval32(P, o, d);
            const float span = iv.upper - iv.lower;
            const float seed = seeds[ray];
            if (!isfinite(seed)) flag(P, view, DPT_STATUS_RANGE);
            if (span > 0.0f && seed != 0.0f) {
                active = true;
                step = span / (float)P.samples;
                volume_factor = seed * (len * step);
                const bool finite = finite_span(o, d, iv.lower, step, 0, P.samples);
                bad = !finite;
                const float m = fabsf(volume_factor);
                summed = finite && m >= 0x1p-102f && m <= 0x1p96f;
                scattered = !summed;
            }
        }
    }
    for (int i = threadIdx.x; i < P.brick_cap; i += kBlock) brick[i] = 0.0;

    // ---- slabs of the workgroup's in-loop rays over [first, last)
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float ad[3] = {fabsf(d[0]), fabsf(d[1]), fabsf(d[2])};
    const int mine = ad[0] >= ad[1] && ad[0] >= ad[2] ? 0 : (ad[1] >= ad[2] ? 1 : 2);
    int my_votes[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) my_votes[a] = __popcll(__builtin_amdgcn_ballot_w64(summed && mine == a));
    if (lane == 0)
        for (int a = 0; a < 3; ++a) bound_red[wv][a] = my_votes[a];
    __syncthreads();
    int votes[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) votes[a] = bound_red[0][a] + bound_red[1][a] + bound_red[2][a] + bound_red[3][a];
    const int axis = votes[0] >= votes[1] && votes[0] >= votes[2] ? 0 : (votes[1] >= votes[2] ? 1 : 2);
    const int n_axis = axis == 0 ? P.nx : (axis == 1 ? P.ny : P.nz);
    const float oa = axis == 0 ? o[0] : (axis == 1 ? o[1] : o[2]);
    const float da = axis == 0 ? d[0] : (axis == 1 ? d[1] : d[2]);
    int lo = INT_MAX, nhi = INT_MAX, nadv = INT_MAX;
    if (summed) {
        const float q0 = fmaf(fmaf((float)first + 0.5f, step, iv.lower), da, oa);
        const float q1 = fmaf(fmaf((float)(last - 1) + 0.5f, step, iv.lower), da, oa);
        lo = (int)floorf(fminf(fmaxf(fminf(q0, q1), -2.0f), (float)n_axis + 1.0f));
        nhi = -(int)ceilf(fminf(fmaxf(fmaxf(q0, q1), -2.0f), (float)n_axis + 1.0f));
        // advance per sample in 1/64 grid samples, at least 1
        nadv = -(int)fminf(fmaxf(ceilf(step * fabsf(da) * 64.0f), 1.0f), 1048576.0f);
    }
    lo = wave_min(lo);
    nhi = wave_min(nhi);
    nadv = wave_min(nadv);
    __syncthreads();  // every wave has read the votes
    if (lane == 0) {
        bound_red[wv][3] = lo;
        bound_red[wv][4] = nhi;
        bound_red[wv][5] = nadv;
    }
    __syncthreads();
    int slab_lo = min(min(bound_red[0][3], bound_red[1][3]), min(bound_red[2][3], bound_red[3][3]));
    const int slab_hi = -min(min(bound_red[0][4], bound_red[1][4]), min(bound_red[2][4], bound_red[3][4]));
    const int adv = -min(min(bound_red[0][5], bound_red[1][5]), min(bound_red[2][5], bound_red[3][5]));
    __syncthreads();  // bound_red is reused by the chunk bounds
    int width = 1, count = 1;
    if (slab_lo == INT_MAX) {
        slab_lo = 0;
    } else {
        width = max(1, (P.chunk_samples * adv + 63) / 64);
        count = max(1, (slab_hi - slab_lo + width - 1) / width);
    }

    CellSums cs;
    cs.reset();
    for (int ch = 0; ch < count; ++ch) {
        // this lane's samples [k0, k1) in slab ch
        int k0 = first, k1 = first;
        if (summed) {
            if (da == 0.0f) {
                const int own = min(max((int)floorf((oa - (float)slab_lo) / (float)width), 0), count - 1);
                k1 = ch == own ? last : first;
            } else {
                // t(b): the sample index where the coordinate crosses b
                const int js = da > 0.0f ? ch : count - 1 - ch;
                const float b0 = (float)(slab_lo + js * width), b1 = (float)(slab_lo + (js + 1) * width);
                const float t0 = ((b0 - oa) / da - iv.lower) / step - 0.5f;
                const float t1 = ((b1 - oa) / da - iv.lower) / step - 0.5f;
                const float lo_k = (float)first, hi_k = (float)last;
                if (da > 0.0f) {
                    k0 = ch == 0 ? first : (int)fminf(fmaxf(ceilf(t0), lo_k), hi_k);
                    k1 = ch == count - 1 ? last : (int)fminf(fmaxf(ceilf(t1), lo_k), hi_k);
                } else {
                    k0 = ch == 0 ? first : (int)fminf(fmaxf(floorf(t1) + 1.0f, lo_k), hi_k);
                    k1 = ch == count - 1 ? last : (int)fminf(fmaxf(floorf(t0) + 1.0f, lo_k), hi_k);
                }
                k1 = max(k1, k0);
            }
        }
        int blo[3] = {INT_MAX, INT_MAX, INT_MAX}, bnhi[3] = {INT_MAX, INT_MAX, INT_MAX};
        if (k1 > k0) {
            const float pa = fmaf((float)k0 + 0.5f, step, iv.lower);
            const float pb = fmaf((float)(k1 - 1) + 0.5f, step, iv.lower);
            int la[3], lb[3];
            lower_cell32(P, fmaf(pa, d[0], o[0]), fmaf(pa, d[1], o[1]), fmaf(pa, d[2], o[2]), la);
            lower_cell32(P, fmaf(pb, d[0], o[0]), fmaf(pb, d[1], o[1]), fmaf(pb, d[2], o[2]), lb);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                blo[a] = min(la[a], lb[a]);
                bnhi[a] = -max(la[a], lb[a]);
            }
        }
        Brick B{};
        const bool use_lds = chunk_brick(P, bound_red, blo, bnhi, P.nx > 1, B, nullptr, true);
        int cur = -1;
        for (int k = k0; k < k1; ++k) {
            const float pos = fmaf((float)k + 0.5f, step, iv.lower);
            const float p0 = fmaf(pos, d[0], o[0]);
            const float p1 = fmaf(pos, d[1], o[1]);
            const float p2 = fmaf(pos, d[2], o[2]);
            Cell32 c;
            cell32_inside(P, p0, p1, p2, c);
            const bool inside = inside32(P, p0, p1, p2);
            // an outside sample adds nothing (zero fractions and factor)
            c.fx = inside ? c.fx : 0.0f;
            c.fy = inside ? c.fy : 0.0f;
            c.fz = inside ? c.fz : 0.0f;
            const float factor = inside ? volume_factor : 0.0f;
            if (c.base != cur) {
                if (cs.x >= 0) {
                    if (use_lds)
                        flush_lds(B, cs, brick);
                    else
                        flush_device(P, cs, acc);
                }
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) cs.s[kk] = 0.0f;
                cs.x = c.lx;
                cs.y = c.ly;
                cs.z = c.lz;
                cur = c.base;
            }
            float wt[8];
            weights32(c, wt);
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) cs.s[kk] = fmaf(wt[kk], factor, cs.s[kk]);
        }
        if (cs.x >= 0) {
            if (use_lds)
                flush_lds(B, cs, brick);
            else
                flush_device(P, cs, acc);
        }
        cs.reset();
        __syncthreads();
        if (use_lds) drain_brick(P, B, brick, acc);
        __syncthreads();
    }
    // rays outside the in-loop factor range: per-increment device atomics,
    // as vjp32_kernel
    if (scattered) {
        for (int k = first; k < last; ++k) {
            const float pos = fmaf((float)k + 0.5f, step, iv.lower);
            const float p0 = fmaf(pos, d[0], o[0]);
            const float p1 = fmaf(pos, d[1], o[1]);
            const float p2 = fmaf(pos, d[2], o[2]);
            if (!inside32(P, p0, p1, p2)) continue;
            Cell32 c;
            cell32_inside(P, p0, p1, p2, c);
            float wt[8];
            weights32(c, wt);
            const int voxel = (c.lz * P.ny + c.ly) * P.nx + c.lx;
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) {
                const float inc = P.volume_f64 ? volume_factor * wt[kk] : scatter_increment(volume_factor, wt[kk]);
                bad |= !isfinite(inc);
                if (inc != 0.0f) unsafeAtomicAdd(&acc[voxel + corner_offset(P, kk)], (double)inc);
            }
        }
    }
    if (bad) flag(P, view, DPT_STATUS_RANGE);
}

// --------------------------------------------------------- FP64 sampling --

struct Cell64 {
    int base;
    int lx, ly, lz;
    double fx, fy, fz;
    double slx, sly, slz;
};

// kernels/projection.py:124-134 + 199-214
__device__ __forceinline__ bool cell64(const Params &P, const double *p, Cell64 &c) {
    const double hx = (double)P.nx - 0.5, hy = (double)P.ny - 0.5, hz = (double)P.nz - 0.5;
    if (!(p[0] >= -0.5 && p[1] >= -0.5 && p[2] >= -0.5 && p[0] <= hx && p[1] <= hy && p[2] <= hz))
        return false;
    const double mx = (double)(P.nx - 1), my = (double)(P.ny - 1), mz = (double)(P.nz - 1);
    const double v0 = fmin(fmax(p[0], 0.0), mx);
    const double v1 = fmin(fmax(p[1], 0.0), my);
    const double v2 = fmin(fmax(p[2], 0.0), mz);
    const int lx = min((int)floor(v0), P.capx);
    const int ly = min((int)floor(v1), P.capy);
    const int lz = min((int)floor(v2), P.capz);
    c.lx = lx;
    c.ly = ly;
    c.lz = lz;
    c.fx = v0 - (double)lx;
    c.fy = v1 - (double)ly;
    c.fz = v2 - (double)lz;
    c.slx = (p[0] >= 0.0 && p[0] < mx) ? 1.0 : 0.0;
    c.sly = (p[1] >= 0.0 && p[1] < my) ? 1.0 : 0.0;
    c.slz = (p[2] >= 0.0 && p[2] < mz) ? 1.0 : 0.0;
    c.base = (lz * P.ny + ly) * P.nx + lx;
    return true;
}

// Clamped lower cell of an FP64 sample point (cell64's arithmetic without the
// support test: out-of-support samples only widen a chunk's box).
__device__ __forceinline__ void lower_cell64(const Params &P, const double *p, int *l) {
    l[0] = min((int)floor(fmin(fmax(p[0], 0.0), (double)(P.nx - 1))), P.capx);
    l[1] = min((int)floor(fmin(fmax(p[1], 0.0), (double)(P.ny - 1))), P.capy);
    l[2] = min((int)floor(fmin(fmax(p[2], 0.0), (double)(P.nz - 1))), P.capz);
}

// kernels/projection.py:236-274 (weighted lerps; differences before interpolation)
template <bool GRAD>
__device__ __forceinline__ double sample_cell64(const Cell64 &c, const double *v, double *g) {
    const double v000 = v[0], v100 = v[1], v010 = v[2], v110 = v[3];
    const double v001 = v[4], v101 = v[5], v011 = v[6], v111 = v[7];
    const double fx = c.fx, fy = c.fy, fz = c.fz;
    const double ax = 1.0 - fx, ay = 1.0 - fy, az = 1.0 - fz;
    const double x00 = ax * v000 + fx * v100, x10 = ax * v010 + fx * v110;
    const double x01 = ax * v001 + fx * v101, x11 = ax * v011 + fx * v111;
    const double xy0 = ay * x00 + fy * x10, xy1 = ay * x01 + fy * x11;
    const double value = az * xy0 + fz * xy1;
    if constexpr (GRAD) {
        const double dx00 = v100 - v000, dx10 = v110 - v010, dx01 = v101 - v001, dx11 = v111 - v011;
        const double dx0 = ay * dx00 + fy * dx10, dx1 = ay * dx01 + fy * dx11;
        const double dy0 = ax * (v010 - v000) + fx * (v110 - v100);
        const double dy1 = ax * (v011 - v001) + fx * (v111 - v101);
        const double dz0 = ax * (v001 - v000) + fx * (v101 - v100);
        const double dz1 = ax * (v011 - v010) + fx * (v111 - v110);
        g[0] = c.slx * (az * dx0 + fz * dx1);
        g[1] = c.sly * (az * dy0 + fz * dy1);
        g[2] = c.slz * (ay * dz0 + fy * dz1);
    }
    return value;
}

// The eight corners as FP64, fetched with the x-pair loads of the FP32 path
// on grids with nx > 1 (uniform branch).
template <typename T>
__device__ __forceinline__ void load8d(const Params &P, const T *field, int base, double *v) {
    float f[8];
    if (P.sx)
        load8<T, kLayoutPair>(P, field, base, f);
    else
        load8<T, kLayoutScalar>(P, field, base, f);
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = (double)f[k];
}

// kernels/projection.py:802-846 (midpoint, FP64 sampling)
template <typename T>
__global__ __launch_bounds__(kBlock) void forward64_kernel(Params P, const T *__restrict__ field,
                                                           double *__restrict__ out) {
    int view, row, col, wslot;
    bool wave_valid;
    if (!lane_ray(P, view, row, col, wslot, wave_valid)) return;
    if (!ray_in_range(P, view, row, col)) return;
    const size_t ray = ((size_t)view * P.rows + row) * P.cols + col;
    const double *pose = P.poses + 12 * (size_t)view;
    World w;
    ray_world(P, pose, view, row, col, w);
    Ray r;
    ray_grid(P, w, r);
    if (!r.valid) {
        flag(P, view, DPT_STATUS_RANGE);
        out[ray] = 0.0;
        return;
    }
    const Interval64 iv = interval64(P, r.o, r.d);
    double integral = 0.0;
    if (iv.upper > iv.lower) {
        const double step = (iv.upper - iv.lower) / (double)P.samples;
        bool bad = false;
        double v[8];
        int cur = -1;  // corners cached while consecutive samples share a cell
        for (int k = 0; k < P.samples; ++k) {
            const double pos = iv.lower + ((double)k + 0.5) * step;
            const double p[3] = {r.o[0] + pos * r.d[0], r.o[1] + pos * r.d[1], r.o[2] + pos * r.d[2]};
            bad |= !finite3d(p);
            Cell64 c;
            if (cell64(P, p, c)) {
                if (c.base != cur) {
                    load8d<T>(P, field, c.base, v);
                    cur = c.base;
                }
                integral += sample_cell64<false>(c, v, nullptr);
            }
        }
        if (bad) flag(P, view, DPT_STATUS_RANGE);
        integral *= r.len * step;
    }
    out[ray] = integral;
    if (!isfinite(integral)) flag(P, view, DPT_STATUS_RANGE);
}

// kernels/projection.py:873-1237 (midpoint, FP64 sampling, FP32 increments)
// Slab chunks of the FP64 volume adjoint.  Chunks that follow the sample
// index put the k-th samples of a tile's rays in one brick, which spreads
// apart where the tile's clipped intervals differ in length (at the grid's
// silhouette) and overflows the brick.  Here the workgroup cuts space into
// slabs of W grid samples along its dominant axis A (the axis most of its
// rays advance fastest along) and each lane visits, per chunk, the samples
// whose A coordinate lies in the chunk's slab; W is the advance of the
// fastest lane over the plan's chunk length.  Slab boundaries are integers
// and both sides of a boundary use the same expression, so every sample
// falls in exactly one chunk, and a lane's chunks come in its sample order.
struct Slabs {
    int axis, lo, width, count;
};

__device__ __forceinline__ double pick3(const double *v, int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); }

__device__ __forceinline__ Slabs slab_setup(const Params &P, bool active, const Ray &r, const Interval64 &iv,
                                            double step, int (*red)[6], int from = 0) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double ad[3] = {fabs(r.d[0]), fabs(r.d[1]), fabs(r.d[2])};
    const int mine = ad[0] >= ad[1] && ad[0] >= ad[2] ? 0 : (ad[1] >= ad[2] ? 1 : 2);
    int votes[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) votes[a] = __popcll(__builtin_amdgcn_ballot_w64(active && mine == a));
    if (lane == 0)
        for (int a = 0; a < 3; ++a) red[wv][a] = votes[a];
    __syncthreads();
    int total[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) total[a] = red[0][a] + red[1][a] + red[2][a] + red[3][a];
    const int axis = total[0] >= total[1] && total[0] >= total[2] ? 0 : (total[1] >= total[2] ? 1 : 2);
    const int n = axis == 0 ? P.nx : (axis == 1 ? P.ny : P.nz);
    int lo = INT_MAX, nhi = INT_MAX, nadv = INT_MAX;
    if (active) {
        const double oa = pick3(r.o, axis), da = pick3(r.d, axis);
        const double q0 = oa + (iv.lower + ((double)from + 0.5) * step) * da;
        const double q1 = oa + (iv.lower + ((double)P.samples - 0.5) * step) * da;
        lo = (int)floor(fmin(fmax(fmin(q0, q1), -2.0), (double)n + 1.0));
        nhi = -(int)ceil(fmin(fmax(fmax(q0, q1), -2.0), (double)n + 1.0));
        // advance per sample in 1/64 grid samples, at least 1
        nadv = -(int)fmin(fmax(ceil(step * fabs(da) * 64.0), 1.0), 1048576.0);
    }
    lo = wave_min(lo);
    nhi = wave_min(nhi);
    nadv = wave_min(nadv);
    __syncthreads();  // every wave has read the votes
    if (lane == 0) {
        red[wv][3] = lo;
        red[wv][4] = nhi;
        red[wv][5] = nadv;
    }
    __syncthreads();
    Slabs out;
    out.axis = axis;
    out.lo = min(min(red[0][3], red[1][3]), min(red[2][3], red[3][3]));
    const int hi = -min(min(red[0][4], red[1][4]), min(red[2][4], red[3][4]));
    const int adv = -min(min(red[0][5], red[1][5]), min(red[2][5], red[3][5]));
    __syncthreads();  // red is reused by the chunk bounds
    if (out.lo == INT_MAX) {
        out.lo = 0;
        out.width = 1;
        out.count = 1;
        return out;
    }
    out.width = max(1, (P.chunk_samples * adv + 63) / 64);
    out.count = max(1, (hi - out.lo + out.width - 1) / out.width);
    return out;
}

// Sample range [k0, k1) of this lane in chunk j, over samples [from, S).
__device__ __forceinline__ void slab_range(const Params &P, bool active, const Ray &r, const Interval64 &iv,
                                           double step, const Slabs &sl, int j, int &k0, int &k1, int from = 0) {
    k0 = k1 = from;
    if (!active) return;
    const double oa = pick3(r.o, sl.axis), da = pick3(r.d, sl.axis);
    const double S = (double)P.samples, F = (double)from;
    if (da == 0.0) {
        const int own = min(max((int)floor((oa - (double)sl.lo) / (double)sl.width), 0), sl.count - 1);
        k1 = j == own ? P.samples : from;
        return;
    }
    // t(b) = ((b - o) / d - lower) / step - 0.5, monotone in b: the sample
    // index where the coordinate crosses the boundary b
    const int js = da > 0.0 ? j : sl.count - 1 - j;
    const double b0 = (double)(sl.lo + js * sl.width), b1 = (double)(sl.lo + (js + 1) * sl.width);
    const double t0 = ((b0 - oa) / da - iv.lower) / step - 0.5;
    const double t1 = ((b1 - oa) / da - iv.lower) / step - 0.5;
    if (da > 0.0) {
        // coordinate(k) >= b  <=>  k >= t(b)
        k0 = j == 0 ? from : (int)fmin(fmax(ceil(t0), F), S);
        k1 = j == sl.count - 1 ? P.samples : (int)fmin(fmax(ceil(t1), F), S);
    } else {
        // coordinate(k) < b  <=>  k > t(b)
        k0 = j == 0 ? from : (int)fmin(fmax(floor(t1) + 1.0, F), S);
        k1 = j == sl.count - 1 ? P.samples : (int)fmin(fmax(floor(t0) + 1.0, F), S);
    }
    k1 = max(k1, k0);
}

// The joint form is held to three waves per SIMD (168 VGPRs, about 100 bytes
// of spills, mostly the slab fallback's state): left free it takes 189 and
// two waves, and ran the full scan's FP64 joint VJP in 736 against 590 ms.
template <typename T, bool VOL, bool POSE, int C, bool SLABS = false>
__global__ __launch_bounds__(kBlock, (VOL && POSE) ? 3 : 1) void vjp64_kernel(Params P, const T *__restrict__ field,
                                                       const double *__restrict__ seeds,
                                                       double *__restrict__ acc,
                                                       double *__restrict__ partials) {
    extern __shared__ double brick[];
    __shared__ int bound_red[kBlock / 64][6];
    int view, row, col, wslot;
    bool wave_valid;
    if (!lane_ray(P, view, row, col, wslot, wave_valid)) return;
    double gradient[kMaxComponents];
#pragma unroll
    for (int c = 0; c < kMaxComponents; ++c) gradient[c] = 0.0;
    const double *pose = P.poses + 12 * (size_t)view;

    // ---- per-ray setup
    bool active = false;
    Ray r{};
    Interval64 iv{};
    double span = 0.0, seed = 0.0, factor = 0.0, step = 0.0, volume_factor = 0.0;
    const double S = (double)P.samples;
    if (ray_in_range(P, view, row, col)) {
        const size_t ray = ((size_t)view * P.rows + row) * P.cols + col;
        World w;
        ray_world(P, pose, view, row, col, w);
        ray_grid(P, w, r);
        if (!r.valid) {
            flag(P, view, DPT_STATUS_RANGE);
        } else {
            iv = interval64(P, r.o, r.d);
            span = iv.upper - iv.lower;
            seed = seeds[ray];
            if (seed == 0.0) span = 0.0;
            if (span > 0.0) {
                active = true;
                factor = seed * r.len;
                step = span / S;
                volume_factor = factor * span / S;
                if (!isfinite(factor)) volume_factor = product4(seed, r.len, span, 1.0 / S);
            }
        }
    }
    const bool pow2 = (P.samples & (P.samples - 1)) == 0;
    const double inverse_samples = pow2 ? 1.0 / S : 0.0;
    double mean = 0.0, dor[3] = {0, 0, 0}, ddir[3] = {0, 0, 0};
    double shift_lower = 0.0, shift_upper = 0.0;
    bool bad = false;

    // Lockstep chunks with LDS bricks, as vjp32_kernel: the reference adds
    // every FP32-rounded contribution with its own atomic; here they go to
    // the workgroup's FP64 LDS brick (exact sums of FP32 values) and reach
    // the FP64 accumulator once per touched grid sample and chunk.
    if constexpr (VOL) {
        for (int i = threadIdx.x; i < P.brick_cap; i += kBlock) brick[i] = 0.0;
    }
    double v[8];
    int cur = -1;  // corners cached while consecutive samples share a cell
    // SLABS (volume adjoint only, chosen per plan): chunks by slabs of space.
    // Index chunks switch to slabs for the rest of the workgroup's samples
    // when one overflows the brick (P.slab_fallback); each lane's slabs come
    // in its sample order, so its sums keep their order.
    Slabs slabs{0, 0, 1, 1};
    bool slab_mode = VOL && SLABS;
    int from = 0;
    if constexpr (VOL && SLABS) slabs = slab_setup(P, active, r, iv, step, bound_red);
    int chunks = slab_mode ? slabs.count : (P.samples + P.chunk_samples - 1) / P.chunk_samples;
    for (int ch = 0; ch < chunks; ++ch) {
        int k0, k1;
        if (slab_mode) {
            slab_range(P, active, r, iv, step, slabs, ch, k0, k1, from);
        } else {
            k0 = ch * P.chunk_samples;
            k1 = min(k0 + P.chunk_samples, P.samples);
        }
        Brick B{};
        bool use_lds = false;
        if constexpr (VOL) {
            int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, nhi[3] = {INT_MAX, INT_MAX, INT_MAX};
            if (active && k1 > k0) {
                // sample positions are monotone in k: the chunk's end samples bound it
                const double pa = iv.lower + ((double)k0 + 0.5) * step;
                const double pb = iv.lower + ((double)(k1 - 1) + 0.5) * step;
                const double qa[3] = {r.o[0] + pa * r.d[0], r.o[1] + pa * r.d[1], r.o[2] + pa * r.d[2]};
                const double qb[3] = {r.o[0] + pb * r.d[0], r.o[1] + pb * r.d[1], r.o[2] + pb * r.d[2]};
                int la[3], lb[3];
                lower_cell64(P, qa, la);
                lower_cell64(P, qb, lb);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    lo[a] = min(la[a], lb[a]);
                    nhi[a] = -max(la[a], lb[a]);
                }
            }
            bool overflow = false;
            use_lds = chunk_brick(P, bound_red, lo, nhi, false, B, &overflow);
            if (overflow && !slab_mode && P.slab_fallback) {  // uniform
                __syncthreads();  // the chunk bound is read
                from = k0;
                slab_mode = true;
                slabs = slab_setup(P, active, r, iv, step, bound_red, from);
                chunks = slabs.count;
                ch = -1;
                continue;
            }
        }
        if (active) {
            for (int k = k0; k < k1; ++k) {
                const double alpha = pow2 ? ((double)k + 0.5) * inverse_samples : ((double)k + 0.5) / S;
                const double pos = iv.lower + ((double)k + 0.5) * step;
                const double p[3] = {r.o[0] + pos * r.d[0], r.o[1] + pos * r.d[1], r.o[2] + pos * r.d[2]};
                bad |= !finite3d(p);
                Cell64 c;
                if (!cell64(P, p, c)) continue;
                if constexpr (POSE) {
                    double g[3];
                    if (c.base != cur) {
                        load8d<T>(P, field, c.base, v);
                        cur = c.base;
                    }
                    const double value = sample_cell64<true>(c, v, g);
                    mean += value;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        dor[a] += g[a];
                        ddir[a] += pos * g[a];
                    }
                    const double along = g[0] * r.d[0] + g[1] * r.d[1] + g[2] * r.d[2];
                    shift_lower += (1.0 - alpha) * along;
                    shift_upper += alpha * along;
                }
                if constexpr (VOL) {
                    // kernels/projection.py:396-430: seed * weights in axis order, one FP32 rounding
                    const int local = use_lds ? ((c.lz - B.z0) * B.by + (c.ly - B.y0)) * B.bx + (c.lx - B.x0) : 0;
#pragma unroll
                    for (int kk = 0; kk < 8; ++kk) {
                        double contribution = volume_factor;
                        contribution *= (kk & 1) ? c.fx : 1.0 - c.fx;
                        contribution *= (kk & 2) ? c.fy : 1.0 - c.fy;
                        contribution *= (kk & 4) ? c.fz : 1.0 - c.fz;
                        const double rounded = destination_increment(P, contribution);
                        if (!isfinite(rounded)) bad = true;
                        if (rounded == 0.0) continue;
                        if (use_lds) {
                            const int off = ((kk & 1) ? B.lsx : 0) + ((kk & 2) ? B.lsy : 0) + ((kk & 4) ? B.lsz : 0);
                            atomicAdd(&brick[local + off], rounded);
                        } else {
                            unsafeAtomicAdd(&acc[c.base + corner_offset(P, kk)], rounded);
                        }
                    }
                }
            }
        }
        if constexpr (VOL) {
            __syncthreads();
            if (use_lds) drain_brick(P, B, brick, acc);
            __syncthreads();
        }
    }
    if (bad) flag(P, view, DPT_STATUS_RANGE);
    if constexpr (POSE) {
        if (active) {
            const double lower_seed = -mean + span * shift_lower;
            const double upper_seed = mean + span * shift_upper;
            double og[3], dg[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double lo_o = (iv.lo_axis == a) ? iv.lo_o : 0.0;
                const double lo_d = (iv.lo_axis == a) ? iv.lo_d : 0.0;
                const double up_o = (iv.up_axis == a) ? iv.up_o : 0.0;
                const double up_d = (iv.up_axis == a) ? iv.up_d : 0.0;
                og[a] = span * dor[a] + lower_seed * lo_o + upper_seed * up_o;
                dg[a] = span * ddir[a] + lower_seed * lo_d + upper_seed * up_d;
            }
            World w;  // recomputed: keeps its FP64 registers out of the sample loop
            ray_world(P, pose, view, row, col, w);
            const double f = factor / S;
            if (isfinite(f)) {
                pose_epilogue<C>(P, pose, w, f, og, dg, gradient, seed, mean * step, true);
            } else {
                // kernels/projection.py:1163-1183 range-safe products
                double sa[3], da[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const double s = P.g[i] * og[0] + P.g[3 + i] * og[1] + P.g[6 + i] * og[2];
                    const double t = P.g[i] * dg[0] + P.g[3 + i] * dg[1] + P.g[6 + i] * dg[2];
                    sa[i] = product4(seed, r.len, 1.0 / S, s);
                    da[i] = product4(seed, r.len, 1.0 / S, t);
                }
                Params unit = P;
                for (int i = 0; i < 9; ++i) unit.g[i] = (i % 4 == 0) ? 1.0 : 0.0;
                pose_epilogue<C>(unit, pose, w, 1.0, sa, da, gradient, seed, mean * step, true);
            }
        }
        store_gradient<C>(P, gradient, partials, wslot, wave_valid, view, row, col);
    }
}

// ------------------------------------------------ exact cell-Gauss (FP64) --
// kernels/projection.py:522-746: the clipped ray is cut at every
// interpolation plane it crosses and each piece gets two Gauss nodes, exact
// for the cubic a trilinear field is along a line inside one cell.  Same
// operation order as the oracle, so images agree bit for bit.

struct CellWalk {
    int plane[3], step[3];
    double next[3], inv[3];
};

struct CellPiece {
    double end, radius, first_position, second_position;
    double first[3], second[3];
    int lower[3];
    bool shared;
};

// kernels/projection.py:629-640: cached reciprocal with one FMA residual
// correction; range extremes keep the division
__device__ __forceinline__ double crossing_quotient(double delta, double direction, double inverse) {
    if (inverse == 0.0) return delta / direction;
    const double q = delta * inverse;
    if (!isfinite(q) || fabs(q) < 2.2250738585072014e-308) return delta / direction;
    return fma(fma(-q, direction, delta), inverse, q);
}

// kernels/projection.py:643-655
__device__ __forceinline__ double cell_crossing(const Ray &r, int a, int plane, int n, double upper, double inv) {
    if (r.d[a] == 0.0 || plane < 0 || plane >= n) return upper;
    const double c = crossing_quotient((double)plane - r.o[a], r.d[a], inv);
    return c < upper ? c : upper;
}

// kernels/projection.py:659-695 (begin_cells)
__device__ __forceinline__ void begin_cells(const Params &P, const Ray &r, const Interval64 &iv, CellWalk &w) {
    const int n[3] = {P.nx, P.ny, P.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double entry = r.o[a] + iv.lower * r.d[a];
        const double v = r.d[a];
        w.inv[a] = 0.0;
        w.step[a] = 0;
        w.plane[a] = 0;
        if (v != 0.0) {
            const double inverse = 1.0 / v;
            if (isfinite(inverse) && fabs(inverse) >= 2.2250738585072014e-308 && fabs(v) >= 2.2250738585072014e-308)
                w.inv[a] = inverse;
        }
        if (v > 0.0) {
            w.step[a] = 1;
            w.plane[a] = max((int)ceil(entry), 0);
        } else if (v < 0.0) {
            w.step[a] = -1;
            w.plane[a] = min((int)floor(entry), n[a] - 1);
        }
        double crossing = cell_crossing(r, a, w.plane[a], n[a], iv.upper, w.inv[a]);
        if (crossing <= iv.lower) {
            w.plane[a] += w.step[a];
            crossing = cell_crossing(r, a, w.plane[a], n[a], iv.upper, w.inv[a]);
        }
        w.next[a] = crossing;
    }
}

// kernels/projection.py:710-742 (cell_segment)
__device__ __forceinline__ void cell_piece(const Params &P, const Ray &r, const CellWalk &w, double start,
                                           CellPiece &s) {
    const double m = w.next[1] < w.next[2] ? w.next[1] : w.next[2];
    s.end = w.next[0] < m ? w.next[0] : m;
    s.radius = (s.end - start) * 0.5;
    const double centre = start + s.radius;
    const double offset = s.radius / sqrt(3.0);
    s.first_position = centre - offset;
    s.second_position = centre + offset;
    double mid[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s.first[a] = r.o[a] + s.first_position * r.d[a];
        s.second[a] = r.o[a] + s.second_position * r.d[a];
        mid[a] = r.o[a] + centre * r.d[a];
    }
    lower_cell64(P, mid, s.lower);
    const int n[3] = {P.nx, P.ny, P.nz};
    s.shared = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double lo = (double)s.lower[a], up = lo + 1.0, top = (double)(n[a] - 1);
        const double c1 = fmin(fmax(s.first[a], 0.0), top), c2 = fmin(fmax(s.second[a], 0.0), top);
        if (c1 < lo || c2 < lo) s.shared = false;
        if (s.lower[a] < n[a] - 2 && (c1 >= up || c2 >= up)) s.shared = false;
    }
}

// kernels/projection.py:745-755 (advance_cells): every tied plane advances
__device__ __forceinline__ void advance_cells(const Params &P, const Ray &r, CellWalk &w, double end, double upper) {
    const int n[3] = {P.nx, P.ny, P.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (w.next[a] <= end) {
            w.plane[a] += w.step[a];
            w.next[a] = cell_crossing(r, a, w.plane[a], n[a], upper, w.inv[a]);
        }
}

__device__ __forceinline__ bool inside_support64(const Params &P, const double *p) {
    return p[0] >= -0.5 && p[1] >= -0.5 && p[2] >= -0.5 && p[0] <= (double)P.nx - 0.5 &&
           p[1] <= (double)P.ny - 0.5 && p[2] <= (double)P.nz - 0.5;
}

// Value (and slope) of one node: sample_cell in the piece's cell when it is
// shared (kernels/projection.py:278-284), else sample_field (:288-295); zero
// outside the support.
template <typename T, bool GRAD>
__device__ __forceinline__ double node_value(const Params &P, const T *field, const double *pt, const int *lower,
                                             double *g) {
    if (GRAD) g[0] = g[1] = g[2] = 0.0;
    Cell64 c;
    if (!cell64(P, pt, c)) return 0.0;
    if (lower != nullptr) {
        c.lx = lower[0];
        c.ly = lower[1];
        c.lz = lower[2];
        c.fx = fmin(fmax(pt[0], 0.0), (double)(P.nx - 1)) - (double)c.lx;
        c.fy = fmin(fmax(pt[1], 0.0), (double)(P.ny - 1)) - (double)c.ly;
        c.fz = fmin(fmax(pt[2], 0.0), (double)(P.nz - 1)) - (double)c.lz;
        c.base = (c.lz * P.ny + c.ly) * P.nx + c.lx;
    }
    double v[8];
    load8d<T>(P, field, c.base, v);
    return sample_cell64<GRAD>(c, v, g);
}

// Value (and slope) of a node in a shared piece whose 8 corner values v of
// the cell at lower are already loaded (sample_cell, kernels/projection.py:
// 278-284); zero outside the support.
template <bool GRAD>
__device__ __forceinline__ double node_in_cell(const Params &P, const double *pt, const int *lower, const double *v,
                                               double *g) {
    if (GRAD) g[0] = g[1] = g[2] = 0.0;
    Cell64 c;
    if (!cell64(P, pt, c)) return 0.0;
    c.lx = lower[0];
    c.ly = lower[1];
    c.lz = lower[2];
    c.fx = fmin(fmax(pt[0], 0.0), (double)(P.nx - 1)) - (double)c.lx;
    c.fy = fmin(fmax(pt[1], 0.0), (double)(P.ny - 1)) - (double)c.ly;
    c.fz = fmin(fmax(pt[2], 0.0), (double)(P.nz - 1)) - (double)c.lz;
    return sample_cell64<GRAD>(c, v, g);
}

// Grid index of corner k of the cell at lower (kernels/projection.py:217-222)
__device__ __forceinline__ int corner_of(const Params &P, const int *lower, int k) {
    const int x = min(lower[0] + (k & 1), P.nx - 1);
    const int y = min(lower[1] + ((k >> 1) & 1), P.ny - 1);
    const int z = min(lower[2] + ((k >> 2) & 1), P.nz - 1);
    return (z * P.ny + y) * P.nx + x;
}

// kernels/projection.py:802-846 with cell_gauss: integral * length
template <typename T>
__global__ __launch_bounds__(kBlock) void forward_gauss_kernel(Params P, const T *__restrict__ field,
                                                               double *__restrict__ out) {
    int view, row, col, wslot;
    bool wave_valid;
    if (!lane_ray(P, view, row, col, wslot, wave_valid)) return;
    if (!ray_in_range(P, view, row, col)) return;
    const size_t ray = ((size_t)view * P.rows + row) * P.cols + col;
    const double *pose = P.poses + 12 * (size_t)view;
    World w;
    ray_world(P, pose, view, row, col, w);
    Ray r;
    ray_grid(P, w, r);
    if (!r.valid) {
        flag(P, view, DPT_STATUS_RANGE);
        out[ray] = 0.0;
        return;
    }
    const Interval64 iv = interval64(P, r.o, r.d);
    double integral = 0.0;
    if (iv.upper > iv.lower) {
        CellWalk walk;
        begin_cells(P, r, iv, walk);
        double start = iv.lower;
        const long long limit = (long long)P.nx + P.ny + P.nz + 1;
        bool bad = false;
        for (long long segments = 0; start < iv.upper && segments < limit; ++segments) {
            CellPiece s;
            cell_piece(P, r, walk, start, s);
            if (s.end <= start) {
                bad = true;
                break;
            }
            bad |= !finite3d(s.first) || !finite3d(s.second);
            if (s.shared) {  // one load of the cell for both nodes
                double v[8];
                load8d<T>(P, field, (s.lower[2] * P.ny + s.lower[1]) * P.nx + s.lower[0], v);
                integral += s.radius * node_in_cell<false>(P, s.first, s.lower, v, nullptr);
                integral += s.radius * node_in_cell<false>(P, s.second, s.lower, v, nullptr);
            } else {
                integral += s.radius * node_value<T, false>(P, field, s.first, nullptr, nullptr);
                integral += s.radius * node_value<T, false>(P, field, s.second, nullptr, nullptr);
            }
            start = s.end;
            advance_cells(P, r, walk, s.end, iv.upper);
        }
        bad |= start < iv.upper;
        if (bad) flag(P, view, DPT_STATUS_RANGE);
        integral *= r.len;
    }
    out[ray] = integral;
    if (!isfinite(integral)) flag(P, view, DPT_STATUS_RANGE);
}

// Increment of one corner into the FP64 accumulator, rounded as the
// destination requires (zeros add nothing and are skipped)
__device__ __forceinline__ bool add_increment(const Params &P, double *acc, int index, double contribution) {
    const double rounded = destination_increment(P, contribution);
    if (rounded != 0.0) unsafeAtomicAdd(acc + index, rounded);
    return isfinite(rounded);
}

// Destination of the cell-Gauss volume increments during one chunk: the
// workgroup's LDS brick when the chunk's cells fit it, else the FP64 device
// accumulator.  A cell outside the chunk's bound (a crossing rounded to the
// other side of a plane) also goes to the device accumulator.  Only for grids
// whose axes all have >= 2 samples (corner k of a cell is its lower corner
// plus the bits of k).
struct GaussSink {
    Brick B;
    bool lds;

    __device__ __forceinline__ void add(const Params &P, double *brick, double *acc, const int *l, int k,
                                        double v) const {
        const int x = l[0] - B.x0, y = l[1] - B.y0, z = l[2] - B.z0;
        if (lds && (unsigned)x <= (unsigned)(B.bx - 2) && (unsigned)y <= (unsigned)(B.by - 2) &&
            (unsigned)z <= (unsigned)(B.bz - 2)) {
            atomicAdd(brick + ((z + ((k >> 2) & 1)) * B.by + y + ((k >> 1) & 1)) * B.bx + x + (k & 1), v);
        } else {
            unsafeAtomicAdd(acc + corner_of(P, l, k), v);
        }
    }
};

// Each chunk of the cell-Gauss volume adjoint can be halved this many times
// (8 = 2^3 units per chunk) while its brick overflows LDS.
constexpr int kGaussSplit = 8;

// kernels/projection.py:873-1237 with cell_gauss (FP32 volume destination:
// each increment rounded to FP32, summed here in FP64).
//
// The volume adjoint walks a workgroup's rays in lockstep chunks of their
// clip intervals, P.gauss_chunks per ray: a lane takes the pieces that start
// in its ray's chunk, whose cells lie between the cells of the chunk's end
// points (lower cell coordinates are monotone along a line).  The workgroup
// bounds those cells, privatises the box in an FP64 LDS brick when it fits
// (a chunk is halved, down to 1/8, while it overflows), adds each piece's 8
// rounded corner increments to it with LDS atomics and adds the non-zero
// entries to the FP64 accumulator once per touched grid sample and chunk.
// (Carrying the 4 corners consecutive pieces share in registers halves the
// LDS atomics but costs more in divergent shifts than it saves.)
template <typename T, bool VOL, bool POSE, int C>
__global__ __launch_bounds__(kBlock) void vjp_gauss_kernel(Params P, const T *__restrict__ field,
                                                           const double *__restrict__ seeds,
                                                           double *__restrict__ acc,
                                                           double *__restrict__ partials) {
    extern __shared__ double brick[];
    __shared__ int bound_red[kBlock / 64][6];
    int view, row, col, wslot;
    bool wave_valid;
    if (!lane_ray(P, view, row, col, wslot, wave_valid)) return;
    double gradient[kMaxComponents];
#pragma unroll
    for (int c = 0; c < kMaxComponents; ++c) gradient[c] = 0.0;
    const double *pose = P.poses + 12 * (size_t)view;

    // ---- per-ray setup
    bool active = false, bad = false;
    World w{};
    Ray r{};
    Interval64 iv{};
    double seed = 0.0, span = 0.0;
    if (ray_in_range(P, view, row, col)) {
        const size_t ray = ((size_t)view * P.rows + row) * P.cols + col;
        ray_world(P, pose, view, row, col, w);
        ray_grid(P, w, r);
        if (!r.valid) {
            flag(P, view, DPT_STATUS_RANGE);
        } else {
            iv = interval64(P, r.o, r.d);
            span = iv.upper - iv.lower;
            seed = seeds[ray];
            active = seed != 0.0 && span > 0.0;
        }
    }
    const double factor = seed * r.len;
    double og[3] = {0.0, 0.0, 0.0}, dg[3] = {0.0, 0.0, 0.0};
    double integral = 0.0;  // sum of radius x node value (ray coordinates)
    const bool carried = P.nx > 1 && P.ny > 1 && P.nz > 1;
    CellWalk walk{};
    double start = iv.lower;
    bool walking = active;
    if (active) begin_cells(P, r, iv, walk);
    const long long limit = (long long)P.nx + P.ny + P.nz + 1;
    long long segments = 0;

    if constexpr (VOL) {
        for (int i = threadIdx.x; i < P.brick_cap; i += kBlock) brick[i] = 0.0;
    }
    // chunk bounds in units of 1 / units of each ray's clip interval
    const int units = (VOL && carried) ? P.gauss_chunks * kGaussSplit : 1;
    for (int u0 = 0, u1 = 0; u0 < units; u0 = u1) {
        GaussSink sink{};
        u1 = units;
        if (VOL && carried) {
            for (int len = kGaussSplit;; len >>= 1) {
                u1 = min(u0 + len, units);
                int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, nhi[3] = {INT_MAX, INT_MAX, INT_MAX};
                if (walking) {
                    const double ta = iv.lower + span * ((double)u0 / (double)units);
                    const double tb = u1 == units ? iv.upper : iv.lower + span * ((double)u1 / (double)units);
                    const double qa[3] = {r.o[0] + ta * r.d[0], r.o[1] + ta * r.d[1], r.o[2] + ta * r.d[2]};
                    const double qb[3] = {r.o[0] + tb * r.d[0], r.o[1] + tb * r.d[1], r.o[2] + tb * r.d[2]};
                    int la[3], lb[3];
                    lower_cell64(P, qa, la);
                    lower_cell64(P, qb, lb);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = min(la[a], lb[a]);
                        nhi[a] = -max(la[a], lb[a]);
                    }
                }
                bool overflow = false;
                sink.lds = chunk_brick(P, bound_red, lo, nhi, false, sink.B, &overflow);
                if (!overflow || len <= 1) break;
                __syncthreads();  // the next bound rewrites bound_red
            }
        }
        const double stop = u1 == units ? iv.upper : iv.lower + span * ((double)u1 / (double)units);
        if (walking) {
            for (; start < stop && segments < limit; ++segments) {
                CellPiece s;
                cell_piece(P, r, walk, start, s);
                if (s.end <= start) {
                    bad = true;
                    walking = false;
                    break;
                }
                bad |= !finite3d(s.first) || !finite3d(s.second);
                if constexpr (POSE) {
                    double g1[3], g2[3], v1, v2;
                    if (s.shared) {  // one load of the cell for both nodes
                        double v[8];
                        load8d<T>(P, field, (s.lower[2] * P.ny + s.lower[1]) * P.nx + s.lower[0], v);
                        v1 = node_in_cell<true>(P, s.first, s.lower, v, g1);
                        v2 = node_in_cell<true>(P, s.second, s.lower, v, g2);
                    } else {
                        v1 = node_value<T, true>(P, field, s.first, nullptr, g1);
                        v2 = node_value<T, true>(P, field, s.second, nullptr, g2);
                    }
                    if constexpr (C == kRayCoordinates) {
                        integral += s.radius * v1;
                        integral += s.radius * v2;
                    }
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        og[a] += s.radius * g1[a];
                        og[a] += s.radius * g2[a];
                    }
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        dg[a] += s.radius * s.first_position * g1[a];
                        dg[a] += s.radius * s.second_position * g2[a];
                    }
                }
                if constexpr (VOL) {
                    double vf = factor * s.radius;
                    if (!isfinite(factor)) vf = product3(seed, r.len, s.radius);
                    if (s.shared) {
                        // kernels/projection.py:405-430 (scatter_cell)
                        double f1[3], f2[3];
                        const int n[3] = {P.nx, P.ny, P.nz};
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            const double top = (double)(n[a] - 1);
                            f1[a] = fmin(fmax(s.first[a], 0.0), top) - (double)s.lower[a];
                            f2[a] = fmin(fmax(s.second[a], 0.0), top) - (double)s.lower[a];
                        }
                        const double w1 = inside_support64(P, s.first) ? vf : 0.0;
                        const double w2 = inside_support64(P, s.second) ? vf : 0.0;
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            double first = w1, second = w2;
#pragma unroll
                            for (int a = 0; a < 3; ++a) first *= ((k >> a) & 1) ? f1[a] : 1.0 - f1[a];
#pragma unroll
                            for (int a = 0; a < 3; ++a) second *= ((k >> a) & 1) ? f2[a] : 1.0 - f2[a];
                            if (carried) {
                                const double rounded = destination_increment(P, first + second);
                                bad |= !isfinite(rounded);
                                if (rounded != 0.0) sink.add(P, brick, acc, s.lower, k, rounded);
                            } else {
                                bad |= !add_increment(P, acc, corner_of(P, s.lower, k), first + second);
                            }
                        }
                    } else {
                        // kernels/projection.py:437-470 (scatter per node)
#pragma unroll
                        for (int node = 0; node < 2; ++node) {
                            const double *pt = node == 0 ? s.first : s.second;
                            Cell64 c;
                            if (!cell64(P, pt, c)) continue;
                            const int lower[3] = {c.lx, c.ly, c.lz};
                            const double f[3] = {c.fx, c.fy, c.fz};
#pragma unroll
                            for (int k = 0; k < 8; ++k) {
                                double contribution = vf;
#pragma unroll
                                for (int a = 0; a < 3; ++a) contribution *= ((k >> a) & 1) ? f[a] : 1.0 - f[a];
                                bad |= !add_increment(P, acc, corner_of(P, lower, k), contribution);
                            }
                        }
                    }
                }
                start = s.end;
                advance_cells(P, r, walk, s.end, iv.upper);
            }
        }
        if constexpr (VOL) {
            if (sink.lds) {
                __syncthreads();
                drain_brick(P, sink.B, brick, acc);
            }
            __syncthreads();
        }
    }
    if (active) {
        bad |= start < iv.upper;
        if constexpr (POSE) {
            // kernels/projection.py:1031-1046: endpoint terms of the
            // zero-extended support at the clamped clip points
            double lo_pt[3], up_pt[3];
            const int n[3] = {P.nx, P.ny, P.nz};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double hi = (double)n[a] - 0.5;
                lo_pt[a] = fmin(fmax(r.o[a] + iv.lower * r.d[a], -0.5), hi);
                up_pt[a] = fmin(fmax(r.o[a] + iv.upper * r.d[a], -0.5), hi);
            }
            const double lower_value = node_value<T, false>(P, field, lo_pt, nullptr, nullptr);
            const double upper_value = node_value<T, false>(P, field, up_pt, nullptr, nullptr);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double lo_o = (iv.lo_axis == a) ? iv.lo_o : 0.0, lo_d = (iv.lo_axis == a) ? iv.lo_d : 0.0;
                const double up_o = (iv.up_axis == a) ? iv.up_o : 0.0, up_d = (iv.up_axis == a) ? iv.up_d : 0.0;
                og[a] += upper_value * up_o - lower_value * lo_o;
                dg[a] += upper_value * up_d - lower_value * lo_d;
            }
            if (isfinite(factor)) {
                pose_epilogue<C>(P, pose, w, factor, og, dg, gradient, seed, integral, true);
            } else {
                double sa[3], da[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const double sg = P.g[i] * og[0] + P.g[3 + i] * og[1] + P.g[6 + i] * og[2];
                    const double rg = P.g[i] * dg[0] + P.g[3 + i] * dg[1] + P.g[6 + i] * dg[2];
                    sa[i] = product4(seed, r.len, 1.0, sg);
                    da[i] = product4(seed, r.len, 1.0, rg);
                }
                Params unit = P;
                for (int i = 0; i < 9; ++i) unit.g[i] = (i % 4 == 0) ? 1.0 : 0.0;
                pose_epilogue<C>(unit, pose, w, 1.0, sa, da, gradient, seed, integral, true);
            }
        }
    }
    if (bad) flag(P, view, DPT_STATUS_RANGE);
    if constexpr (POSE) store_gradient<C>(P, gradient, partials, wslot, wave_valid, view, row, col);
}

// ------------------------------------------------------- small kernels ----

// Fixed-order FP64 sum of the tile partials of one view (kernels/reductions.py
// role), then write or accumulate and flag non-finite results.
template <int C>
__global__ __launch_bounds__(kBlock) void pose_reduce_kernel(const double *__restrict__ partials,
                                                             int tiles_per_view, double *grad_pose,
                                                             int accumulate, int *status) {
    const int view = blockIdx.x;
    // every component at once: per-thread strided sums, then one pairwise
    // tree per component (the order of a component-by-component pass)
    __shared__ double s[C][kBlock];
    const double *base = partials + (size_t)view * tiles_per_view * C;
    double v[C];
#pragma unroll
    for (int c = 0; c < C; ++c) v[c] = 0.0;
    for (int t = threadIdx.x; t < tiles_per_view; t += kBlock) {
#pragma unroll
        for (int c = 0; c < C; ++c) v[c] += base[(size_t)t * C + c];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) s[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int stride = kBlock / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) {
#pragma unroll
            for (int c = 0; c < C; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + stride];
        }
        __syncthreads();
    }
    if (threadIdx.x < C) {
        const int c = threadIdx.x;
        double result = s[c][0];
        if (accumulate) result += grad_pose[(size_t)view * C + c];
        grad_pose[(size_t)view * C + c] = result;
        if (!isfinite(result)) atomicOr(&status[view], DPT_STATUS_RANGE);
    }
}

// kernels/projection_checks.py:20-24 / :36-46 (physical field: finite, >= 0)
template <typename T>
__global__ __launch_bounds__(kBlock) void check_field_kernel(const T *__restrict__ field, int64_t n,
                                                             int *status) {
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const float v = load_value<T>(field, (int)i);
        bad |= !(isfinite(v) && v >= 0.0f);
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&status[0], DPT_STATUS_DOMAIN);
}

// kernels/projection_checks.py:49-58
template <typename T>
__global__ __launch_bounds__(kBlock) void check_finite_kernel(const T *__restrict__ x, int64_t n,
                                                              int *status) {
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        bad |= !isfinite(x[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&status[0], DPT_STATUS_RANGE);
}

// kernels/batched_geometry.py:71-88 + kernels/geometry.py:55-74
__global__ void check_poses_kernel(const double *__restrict__ poses, int views, int *status) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= views) return;
    const bool valid = pose_is_rigid(poses + 12 * (size_t)v);
    if (!valid) atomicOr(&status[v], DPT_STATUS_POSE);
}

__global__ __launch_bounds__(kBlock) void fill_f64_kernel(double *__restrict__ x, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        x[i] = 0.0;
}

__global__ void fill_i32_kernel(int *__restrict__ x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0;
}

// FP64 accumulator -> FP32 or FP64 gradient (write or accumulate) + finite
// check (kernels/projection_checks.py:49-58 applied to the final gradient).
template <typename G>
__global__ __launch_bounds__(kBlock) void finish_volume_kernel(const double *__restrict__ acc,
                                                               G *__restrict__ grad, int64_t n,
                                                               int accumulate, int *status) {
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        // the previous value joins the FP64 sum: one rounding to G, and no
        // FP32 overflow of a sum that the previous value brings back in range
        double sum = acc[i];
        if (accumulate) sum += (double)grad[i];
        const G v = (G)sum;
        grad[i] = v;
        bad |= !isfinite(v);
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&status[0], DPT_STATUS_RANGE);
}

// FP32 -> bf16/f16 with round-to-nearest-even, plus the physical-field check.
template <typename T>
__global__ __launch_bounds__(kBlock) void convert_kernel(const float *__restrict__ src, T *__restrict__ dst,
                                                         int64_t n, int *status) {
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const float v = src[i];
        bad |= !(isfinite(v) && v >= 0.0f);
        if constexpr (std::is_same<T, __half>::value) {
            dst[i] = __float2half_rn(v);
        } else {
            dst[i] = __float2bfloat16(v);
        }
    }
    if (status != nullptr && __any(bad) && (threadIdx.x & 63) == 0) atomicOr(&status[0], DPT_STATUS_DOMAIN);
}

// Linear (nz, ny, nx) field -> cell layout (DPT_STORE_CELLS), one lane per
// cell: the 8 corners of cell (i, j, k) in corner order (bit 0 x, bit 1 y,
// bit 2 z), +1 neighbours clamped to the grid as the gather's zero strides
// do on singleton axes.  E is the element's bit pattern (uint32_t for FP32,
// uint16_t for BF16/FP16); a cell is written with 16-byte stores.
template <typename E>
__global__ __launch_bounds__(kBlock) void cells_kernel(const E *__restrict__ src, E *__restrict__ dst, int nx,
                                                       int ny, int nz, int mx, int my, int cells) {
    for (int c = blockIdx.x * kBlock + threadIdx.x; c < cells; c += gridDim.x * kBlock) {
        const int i = c % mx, r = c / mx;
        const int j = r % my, k = r / my;
        const int i1 = min(i + 1, nx - 1), j1 = min(j + 1, ny - 1), k1 = min(k + 1, nz - 1);
        const int64_t slab = (int64_t)nx * ny;
        const int64_t a = k * slab + (int64_t)j * nx, b = k * slab + (int64_t)j1 * nx;
        const int64_t e = k1 * slab + (int64_t)j * nx, f = k1 * slab + (int64_t)j1 * nx;
        const uint32_t v[8] = {src[a + i], src[a + i1], src[b + i], src[b + i1],
                               src[e + i], src[e + i1], src[f + i], src[f + i1]};
        if constexpr (sizeof(E) == 4) {
            Quad *q = reinterpret_cast<Quad *>(dst) + 2 * (size_t)c;
            q[0] = Quad{v[0], v[1], v[2], v[3]};
            q[1] = Quad{v[4], v[5], v[6], v[7]};
        } else {
            Quad *q = reinterpret_cast<Quad *>(dst) + (size_t)c;
            *q = Quad{v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16), v[6] | (v[7] << 16)};
        }
    }
}

// Count rays whose FP32 clipped interval is non-empty (effective samples =
// count * samples); same setup arithmetic as the forward kernel.
__global__ __launch_bounds__(kBlock) void count_hits_kernel(Params P, unsigned long long *count) {
    int view, row, col, wslot;
    bool wave_valid;
    if (!lane_ray(P, view, row, col, wslot, wave_valid)) return;
    bool hit = false;
    if (ray_in_range(P, view, row, col)) {
        const double *pose = P.poses + 12 * (size_t)view;
        World w;
        ray_world(P, pose, view, row, col, w);
        Ray r;
        ray_grid(P, w, r);
        float o[3] = {(float)r.o[0], (float)r.o[1], (float)r.o[2]};
        float d[3] = {(float)r.d[0], (float)r.d[1], (float)r.d[2]};
        if (r.valid) {
            const Interval32 iv = interval32(P, o, d);
            hit = iv.upper > iv.lower;
        }
    }
    const unsigned long long lanes = __popcll(__ballot(hit));
    if ((threadIdx.x & 63) == 0 && lanes) atomicAdd(count, lanes);
}

// Expand per-view panel records into world segments, one per pixel
// (kernels/batched_geometry.py:9-39): record = source (the beam direction
// for parallel panels), pixel (0,0), u, v, column and row pitch, length
// (parallel source distance, or the cylinder radius), radial vector; mode 0
// flat cone, 1 parallel, 2 cylindrical.
__global__ __launch_bounds__(kBlock) void panel_rays_kernel(const double *__restrict__ meta,
                                                            const int *__restrict__ modes, int64_t per_view,
                                                            int cols, int64_t total, double *__restrict__ src,
                                                            double *__restrict__ end) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock) {
        const int view = (int)(i / per_view);
        const int pixel = (int)(i - (int64_t)view * per_view);
        const int row = pixel / cols, col = pixel - row * cols;
        const double *m = meta + 18 * (size_t)view;
        const double du = m[12], dv = m[13], length = m[14];
        double a[3], b[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = m[k];
            b[k] = m[3 + k] + (double)col * du * m[6 + k] + (double)row * dv * m[9 + k];
        }
        if (modes[view] == 1) {
            for (int k = 0; k < 3; ++k) a[k] = b[k] - length * a[k];
        } else if (modes[view] == 2) {
            const double angle = (double)col * du / length;
            const double c = cos(angle) - 1.0, sn = sin(angle);
            for (int k = 0; k < 3; ++k)
                b[k] = m[3 + k] + c * m[15 + k] + sn * length * m[6 + k] + (double)row * dv * m[9 + k];
        }
        for (int k = 0; k < 3; ++k) {
            src[3 * i + k] = a[k];
            end[3 * i + k] = b[k];
        }
    }
}

}  // namespace dpt

// ================================================================ C ABI ====

using namespace dpt;

struct dpt_plan {
    int device = 0;
    Params base{};
    int arithmetic = 0, storage = 0;
    unsigned needs = 0;
    int64_t voxels = 0, pixels = 0;
    double *d_panels = nullptr, *d_ray_src = nullptr, *d_ray_end = nullptr;
    int *d_status = nullptr;
    int32_t *h_status = nullptr;
    double *d_partials = nullptr;
    int partial_segments = 1;              // d_partials holds this many segments' wave partials
    double *d_acc = nullptr;
    float *d_alpha = nullptr;
    unsigned long long *d_count = nullptr;
    bool cell_gauss = false;  // FP64 exact cell-Gauss integration instead of midpoints
    bool cells = false;       // FP32 sampling gathers from the cell-layout copy
    int mx = 1, my = 1;       // cells along x and y
    int64_t cell_count = 0;
    void *d_cells = nullptr;  // 8 * cell_count elements of the storage type
    int vjp_vb = 1;          // views per workgroup of the VJP launches
    int forward_vb = 1;      // views per workgroup of the forward launch
    bool timing = false;
    double *ray_out = nullptr;             // destination of a ray-coordinate VJP (caller's buffer)
    GatherRay *d_gather_rays = nullptr;    // adjoint gather workspace, allocated at its first call
    GatherView *d_gather_views = nullptr;
    int *d_guard = nullptr;
    double *d_panel_meta = nullptr;  // panel records of a plan made from them (the gather's parallel-panel windows)
    int *d_panel_modes = nullptr;
    int image_cols = 0;              // detector columns of such a plan
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool timed[3] = {false, false, false};
};

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(DPT_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

struct DeviceGuard {
    int previous = -1;
    explicit DeviceGuard(int device) {
        if (hipGetDevice(&previous) != hipSuccess) previous = -1;
        if (previous != device) (void)hipSetDevice(device);
    }
    ~DeviceGuard() {
        int now = -1;
        if (previous >= 0 && hipGetDevice(&now) == hipSuccess && now != previous) (void)hipSetDevice(previous);
    }
};

unsigned grid_blocks(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (unsigned)b;
}

// Tile geometry for `vb` views per workgroup (see lane_ray).
void configure_tiles(Params &P, int vb) {
    P.vb = P.bundle ? 1 : vb;
    if (P.bundle) {
        P.tile_w = kBlock;
        P.tile_h = 1;
        P.wt_x = (P.cols + 63) / 64;
        P.wt_y = 1;
    } else {
        P.tile_w = P.vb == 4 ? 8 : 16;
        P.tile_h = P.vb == 1 ? 16 : 8;
        P.wt_x = (P.cols + 7) / 8;
        P.wt_y = (P.rows + 7) / 8;
    }
    P.wt_per_view = P.wt_x * P.wt_y;
    P.tiles_x = (P.cols + P.tile_w - 1) / P.tile_w;
    P.tiles_y = (P.rows + P.tile_h - 1) / P.tile_h;
    P.tiles_per_group = P.tiles_x * P.tiles_y;
    P.total_tiles = P.tiles_per_group * ((P.views + P.vb - 1) / P.vb);
    P.chunk = (P.total_tiles + kXcds - 1) / kXcds;
    P.segments = 1;
    P.base_tiles = P.total_tiles;
    P.segment_samples = P.samples;
    P.slab_fallback = 1;
    if (const char *fb = std::getenv("DPT_AMD_SLAB_FALLBACK")) P.slab_fallback = std::atoi(fb) != 0;
}

// Sample-range segments of the FP32 VJP: a launch of few tiles (one view of
// 512^2 is 1024 workgroups, one wave per SIMD) leaves every wave walking all
// its samples with its memory latency exposed.  Such launches run each tile
// as `segments` workgroups over consecutive ranges of whole chunks (at least
// two chunks and 256 samples each), aiming at 16384 waves; the pose partials
// of a ray's segments are summed in FP64 by the fixed-order reduction, the
// volume increments of each segment go to the brick as before.  A segmented
// pose gradient is not the reference's one sequential FP32 sum per ray, so
// only 16-bit storage (blocked compensated sums, no FP32 counterpart)
// segments the pose; FP32 storage runs it unsegmented in the pipelined loop.
// DPT_AMD_VJP_SEGMENTS overrides the count (1..8; not the two-chunk floor).
// A launch of at most 8192 waves (8 per SIMD): one whose sample loops expose
// their memory latency.
bool small_launch(const Params &P) { return (long long)P.total_tiles * (kBlock / 64) * 2 <= 16384; }

void configure_segments(Params &P) {
    const long long waves = (long long)P.total_tiles * (kBlock / 64);
    int g = 1;
    while (g < 8 && waves * (2 * g) <= 16384 && P.samples >= 256 * (2 * g)) g *= 2;
    if (const char *forced = std::getenv("DPT_AMD_VJP_SEGMENTS")) {
        const int v = std::atoi(forced);
        if (v >= 1 && v <= 8) g = v;
    }
    const int chunks = (P.samples + P.chunk_samples - 1) / P.chunk_samples;
    g = std::max(1, std::min(g, chunks / 2));
    if (g <= 1) return;
    const int per = (chunks + g - 1) / g;  // whole chunks per segment
    g = (chunks + per - 1) / per;
    if (g <= 1) return;
    P.segments = g;
    P.segment_samples = per * P.chunk_samples;
    P.base_tiles = P.total_tiles;
    P.total_tiles = P.base_tiles * g;
    P.chunk = (P.total_tiles + kXcds - 1) / kXcds;
}

// Views share a workgroup (and its LDS brick) only when their central rays
// differ by at most ~4 degrees: then their samples of one chunk overlap.
int choose_views_per_block(const double *panels, int views, int rows, int cols) {
    if (views < 2) return 1;
    std::vector<double> dir(3 * (size_t)views);
    const int n = views;
    for (int v = 0; v < n; ++v) {
        const double *pn = panels + 12 * (size_t)v;
        double len = 0.0;
        for (int i = 0; i < 3; ++i) {
            const double centre = pn[3 + i] + 0.5 * (cols - 1) * pn[6 + i] + 0.5 * (rows - 1) * pn[9 + i];
            dir[3 * v + i] = centre - pn[i];
            len += dir[3 * v + i] * dir[3 * v + i];
        }
        len = std::sqrt(len);
        for (int i = 0; i < 3; ++i) dir[3 * v + i] /= len;
    }
    const double cos_limit = std::cos(4.0 * 3.14159265358979323846 / 180.0);
    for (int vb : {4, 2}) {
        bool ok = true;
        for (int g = 0; g + 1 < n && ok; g += vb)
            for (int a = g; a < g + vb && a < n && ok; ++a)
                for (int b = a + 1; b < g + vb && b < n; ++b) {
                    const double c = dir[3 * a] * dir[3 * b] + dir[3 * a + 1] * dir[3 * b + 1] + dir[3 * a + 2] * dir[3 * b + 2];
                    if (c < cos_limit) {
                        ok = false;
                        break;
                    }
                }
        if (ok) return vb;
    }
    return 1;
}

Params launch_params(const dpt_plan *p, const double *poses) {
    Params P = p->base;
    P.poses = poses;
    P.status = p->d_status;
    return P;
}

unsigned tile_grid(const Params &P) { return (unsigned)(P.chunk * kXcds); }

// Launch parameters of a gather from the plan's cell-layout copy.
Params cells_params(const dpt_plan *p, Params P) {
    P.ix = p->mx;
    P.iy = p->my;
    return P;
}

template <typename T>
int launch_forward(dpt_plan *p, const Params &P, const void *field, void *out, bool cells, hipStream_t s) {
    const T *f = static_cast<const T *>(field);
    const dim3 grid(tile_grid(P)), block(kBlock);
    if (p->arithmetic == DPT_ARITH_FP32) {
        float *o = static_cast<float *>(out);
        constexpr bool HALF = !std::is_same<T, float>::value;
        if constexpr (!HALF) {
            if (small_launch(P)) {
                if (cells)
                    hipLaunchKernelGGL((forward32_kernel<T, false, kLayoutCells, true>), grid, block, 0, s,
                                       cells_params(p, P), static_cast<const T *>(p->d_cells), o);
                else if (P.nx > 1)
                    hipLaunchKernelGGL((forward32_kernel<T, false, kLayoutPair, true>), grid, block, 0, s, P, f, o);
                else
                    hipLaunchKernelGGL((forward32_kernel<T, false, kLayoutScalar, true>), grid, block, 0, s, P, f,
                                       o);
                HIP_TRY(hipGetLastError());
                return DPT_OK;
            }
        }
        if (cells)
            hipLaunchKernelGGL((forward32_kernel<T, HALF, kLayoutCells>), grid, block, 0, s, cells_params(p, P),
                               static_cast<const T *>(p->d_cells), o);
        else if (P.nx > 1)
            hipLaunchKernelGGL((forward32_kernel<T, HALF, kLayoutPair>), grid, block, 0, s, P, f, o);
        else
            hipLaunchKernelGGL((forward32_kernel<T, HALF, kLayoutScalar>), grid, block, 0, s, P, f, o);
    } else if (p->cell_gauss) {
        hipLaunchKernelGGL((forward_gauss_kernel<T>), grid, block, 0, s, P, f, static_cast<double *>(out));
    } else {
        hipLaunchKernelGGL((forward64_kernel<T>), grid, block, 0, s, P, f, static_cast<double *>(out));
    }
    HIP_TRY(hipGetLastError());
    return DPT_OK;
}

// Pose partials, or the caller's per-ray buffer for ray coordinates and
// per-ray tangents
template <int C>
double *pose_destination(dpt_plan *p) {
    return (C == kRayCoordinates || C == kRayTangent) ? p->ray_out : p->d_partials;
}

template <typename T, int C, int LAYOUT>
void launch_vjp32(dpt_plan *p, const Params &P, const T *f, const float *sd, bool vol, bool pose, bool pipe,
                  hipStream_t s) {
    constexpr bool HALF = !std::is_same<T, float>::value;
    const dim3 grid(tile_grid(P)), block(kBlock);
    const size_t lds = (size_t)P.brick_cap * sizeof(double);
    double *partials = pose_destination<C>(p);
    if constexpr (!HALF && (C == 6 || C == 12)) {
        if (pipe && pose && !vol) {
            hipLaunchKernelGGL((vjp32_kernel<T, false, true, C, false, LAYOUT, true>), grid, block, 0, s, P, f, sd,
                               p->d_acc, partials);
            return;
        }
    }
    if (vol && pose)
        hipLaunchKernelGGL((vjp32_kernel<T, true, true, C, HALF, LAYOUT>), grid, block, lds, s, P, f, sd, p->d_acc,
                           partials);
    else if (pose)
        hipLaunchKernelGGL((vjp32_kernel<T, false, true, C, HALF, LAYOUT>), grid, block, 0, s, P, f, sd, p->d_acc,
                           partials);
    else if constexpr (LAYOUT != kLayoutCells) {  // a volume-only VJP gathers nothing
        if (P.slab_chunks && (C == 6 || C == 12))
            hipLaunchKernelGGL(vol32_slab_kernel, grid, block, lds, s, P, sd, p->d_acc);
        else
            hipLaunchKernelGGL((vjp32_kernel<T, true, false, C, HALF, LAYOUT>), grid, block, lds, s, P, f, sd,
                               p->d_acc, partials);
    }
}

template <typename T, int C>
int launch_vjp_c(dpt_plan *p, const Params &P, const void *field, const void *seeds, bool vol, bool pose,
                 bool cells, hipStream_t s, bool pipe = false) {
    const T *f = static_cast<const T *>(field);
    const dim3 grid(tile_grid(P)), block(kBlock);
    if (p->arithmetic == DPT_ARITH_FP32) {
        const float *sd = static_cast<const float *>(seeds);
        if (cells && pose && P.nx > 1)
            launch_vjp32<T, C, kLayoutCells>(p, cells_params(p, P), static_cast<const T *>(p->d_cells), sd, vol,
                                             pose, pipe, s);
        else if (P.nx > 1)
            launch_vjp32<T, C, kLayoutPair>(p, P, f, sd, vol, pose, pipe, s);
        else
            launch_vjp32<T, C, kLayoutScalar>(p, P, f, sd, vol, pose, pipe, s);
    } else if (p->cell_gauss) {
        const double *sd = static_cast<const double *>(seeds);
        const size_t lds = (size_t)P.brick_cap * sizeof(double);
        double *partials = pose_destination<C>(p);
        if (vol && pose)
            hipLaunchKernelGGL((vjp_gauss_kernel<T, true, true, C>), grid, block, lds, s, P, f, sd, p->d_acc,
                               partials);
        else if (vol)
            hipLaunchKernelGGL((vjp_gauss_kernel<T, true, false, C>), grid, block, lds, s, P, f, sd, p->d_acc,
                               partials);
        else
            hipLaunchKernelGGL((vjp_gauss_kernel<T, false, true, C>), grid, block, 0, s, P, f, sd, p->d_acc,
                               partials);
    } else {
        const double *sd = static_cast<const double *>(seeds);
        const size_t lds = (size_t)P.brick_cap * sizeof(double);
        double *partials = pose_destination<C>(p);
        if (vol && pose && P.slab_chunks)
            hipLaunchKernelGGL((vjp64_kernel<T, true, true, C, true>), grid, block, lds, s, P, f, sd, p->d_acc,
                               partials);
        else if (vol && pose)
            hipLaunchKernelGGL((vjp64_kernel<T, true, true, C>), grid, block, lds, s, P, f, sd, p->d_acc,
                               partials);
        else if (vol && P.slab_chunks)
            hipLaunchKernelGGL((vjp64_kernel<T, true, false, C, true>), grid, block, lds, s, P, f, sd, p->d_acc,
                               partials);
        else if (vol)
            hipLaunchKernelGGL((vjp64_kernel<T, true, false, C>), grid, block, lds, s, P, f, sd, p->d_acc,
                               partials);
        else
            hipLaunchKernelGGL((vjp64_kernel<T, false, true, C>), grid, block, 0, s, P, f, sd, p->d_acc,
                               partials);
    }
    HIP_TRY(hipGetLastError());
    return DPT_OK;
}

template <typename T>
int launch_vjp(dpt_plan *p, const Params &P, const void *field, const void *seeds, bool vol, bool pose,
               int components, bool cells, hipStream_t s, bool pipe = false) {
    if (components == kRayCoordinates) return launch_vjp_c<T, kRayCoordinates>(p, P, field, seeds, false, true, cells, s);
    if (components == kRayTangent) return launch_vjp_c<T, kRayTangent>(p, P, field, seeds, false, true, cells, s);
    if (components == 6) return launch_vjp_c<T, 6>(p, P, field, seeds, vol, pose, cells, s, pipe);
    return launch_vjp_c<T, 12>(p, P, field, seeds, vol, pose, cells, s, pipe);
}

int launch_cells(const void *src, void *dst, int nx, int ny, int nz, int storage, hipStream_t s) {
    const int mx = nx > 1 ? nx - 1 : 1, my = ny > 1 ? ny - 1 : 1, mz = nz > 1 ? nz - 1 : 1;
    const int64_t cells = (int64_t)mx * my * mz;
    if (cells == 0) return DPT_OK;
    if (storage == DPT_STORE_F32)
        hipLaunchKernelGGL(cells_kernel<uint32_t>, dim3(grid_blocks(cells)), dim3(kBlock), 0, s,
                           static_cast<const uint32_t *>(src), static_cast<uint32_t *>(dst), nx, ny, nz, mx, my,
                           (int)cells);
    else
        hipLaunchKernelGGL(cells_kernel<uint16_t>, dim3(grid_blocks(cells)), dim3(kBlock), 0, s,
                           static_cast<const uint16_t *>(src), static_cast<uint16_t *>(dst), nx, ny, nz, mx, my,
                           (int)cells);
    HIP_TRY(hipGetLastError());
    return DPT_OK;
}

int refresh_cells(dpt_plan *p, const void *field, hipStream_t s) {
    return launch_cells(field, p->d_cells, p->base.nx, p->base.ny, p->base.nz, p->storage, s);
}

void timer_begin(dpt_plan *p, int kind, hipStream_t s) {
    if (p->timing) (void)hipEventRecord(p->ev[2 * kind], s);
}
void timer_end(dpt_plan *p, int kind, hipStream_t s) {
    if (p->timing) {
        (void)hipEventRecord(p->ev[2 * kind + 1], s);
        p->timed[kind] = true;
    }
}

}  // namespace

extern "C" {

int32_t dpt_abi_version(void) { return DPT_ABI_VERSION; }

const char *dpt_last_error(void) { return g_error.c_str(); }

int dpt_plan_destroy(dpt_plan *p) {
    if (p == nullptr) return DPT_OK;
    DeviceGuard guard(p->device);
    (void)hipFree(p->d_panels);
    (void)hipFree(p->d_ray_src);
    (void)hipFree(p->d_ray_end);
    (void)hipFree(p->d_status);
    (void)hipFree(p->d_partials);
    (void)hipFree(p->d_acc);
    (void)hipFree(p->d_alpha);
    (void)hipFree(p->d_cells);
    (void)hipFree(p->d_count);
    (void)hipFree(p->d_gather_rays);
    (void)hipFree(p->d_gather_views);
    (void)hipFree(p->d_guard);
    (void)hipFree(p->d_panel_meta);
    (void)hipFree(p->d_panel_modes);
    if (p->h_status) (void)hipHostFree(p->h_status);
    for (auto &e : p->ev)
        if (e) (void)hipEventDestroy(e);
    delete p;
    return DPT_OK;
}

}  // extern "C"

namespace {

// Shared body of dpt_plan_create (panels or host rays) and
// dpt_plan_create_panels (panel records expanded into rays on the device;
// the plan is then a bundle of rows*cols rays per view).
int plan_create(const dpt_grid_desc *grid, int32_t views, int32_t rows, int32_t cols, const double *panels,
                const double *ray_sources, const double *ray_endpoints, const double *panel_meta,
                const int32_t *panel_modes, int32_t samples, int32_t arithmetic, int32_t storage, uint32_t needs,
                int32_t device, dpt_plan **out) {
    if (out == nullptr || grid == nullptr) return fail(DPT_E_INVALID, "grid and out are required");
    const bool generated = panel_meta != nullptr;
    int image_cols = cols;  // columns of one detector row (the bundle below is rows*cols rays per view)
    if (generated) {
        if (panel_modes == nullptr || panels != nullptr || ray_sources != nullptr || ray_endpoints != nullptr)
            return fail(DPT_E_INVALID, "panel records take modes and no panels or rays");
        if (views < 1 || rows < 1 || cols < 1) return fail(DPT_E_INVALID, "views, rows and cols must be positive");
        if ((int64_t)rows * cols > INT32_MAX) return fail(DPT_E_INVALID, "image exceeds signed 32-bit indexing");
        for (int v = 0; v < views; ++v) {
            if (panel_modes[v] < 0 || panel_modes[v] > 2) return fail(DPT_E_INVALID, "panel mode must be 0, 1 or 2");
            for (int k = 0; k < 18; ++k)
                if (!std::isfinite(panel_meta[18 * (size_t)v + k]))
                    return fail(DPT_E_INVALID, "panel records must be finite");
            if (panel_modes[v] != 0 && !(panel_meta[18 * (size_t)v + 14] > 0.0))
                return fail(DPT_E_INVALID, "parallel length / cylinder radius must be positive");
        }
        image_cols = cols;
        cols = rows * cols;
        rows = 1;
    }
    *out = nullptr;
    if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1)
        return fail(DPT_E_INVALID, "grid counts must be positive");
    const int64_t voxels = (int64_t)grid->nx * grid->ny * grid->nz;
    if (voxels > INT32_MAX) return fail(DPT_E_INVALID, "grid exceeds signed 32-bit indexing");
    if ((long long)grid->ny * grid->nz > kMaxSliceProduct)
        return fail(DPT_E_INVALID, "grid ny * nz exceeds 2^24 (24-bit cell index arithmetic)");
    if (views < 1 || rows < 1 || cols < 1) return fail(DPT_E_INVALID, "views, rows and cols must be positive");
    const int64_t pixels = (int64_t)views * rows * cols;
    if (pixels > INT32_MAX) return fail(DPT_E_INVALID, "image exceeds signed 32-bit indexing");
    if (samples < 1) return fail(DPT_E_INVALID, "samples must be positive");
    if (arithmetic != DPT_ARITH_FP32 && arithmetic != DPT_ARITH_FP64)
        return fail(DPT_E_INVALID, "arithmetic must be DPT_ARITH_FP32 or DPT_ARITH_FP64");
    if (storage != DPT_STORE_F32 && storage != DPT_STORE_BF16 && storage != DPT_STORE_F16)
        return fail(DPT_E_INVALID, "storage must be DPT_STORE_F32, DPT_STORE_BF16 or DPT_STORE_F16");
    if (arithmetic == DPT_ARITH_FP32 && (grid->nx > (1 << 24) || grid->ny > (1 << 24) || grid->nz > (1 << 24)))
        return fail(DPT_E_INVALID, "FP32 sampling requires grid axis counts at most 2**24");
    if (!generated && panels == nullptr && (ray_sources == nullptr || ray_endpoints == nullptr))
        return fail(DPT_E_INVALID, "supply panels or explicit ray sources and endpoints");
    if (panels == nullptr && rows != 1) return fail(DPT_E_INVALID, "explicit rays require rows == 1");
    if ((needs & DPT_CELL_GAUSS) && arithmetic != DPT_ARITH_FP64)
        return fail(DPT_E_INVALID, "cell-Gauss integration requires DPT_ARITH_FP64");
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(grid->object_to_grid[i])) return fail(DPT_E_INVALID, "object_to_grid must be finite");

    DeviceGuard guard(device);
    dpt_plan *p = new dpt_plan();
    p->device = device;
    p->arithmetic = arithmetic;
    p->storage = storage;
    p->needs = needs;
    p->cell_gauss = (needs & DPT_CELL_GAUSS) != 0;
    p->voxels = voxels;
    p->pixels = pixels;
    Params &P = p->base;
    P.nx = grid->nx;
    P.ny = grid->ny;
    P.nz = grid->nz;
    P.ix = grid->nx;
    P.iy = grid->ny;
    P.sx = grid->nx > 1 ? 1 : 0;
    P.sy = grid->ny > 1 ? grid->nx : 0;
    P.sz = grid->nz > 1 ? grid->nx * grid->ny : 0;
    P.hx = (float)grid->nx - 0.5f;
    P.hy = (float)grid->ny - 0.5f;
    P.hz = (float)grid->nz - 0.5f;
    P.cx = (float)(grid->nx - 1);
    P.cy = (float)(grid->ny - 1);
    P.cz = (float)(grid->nz - 1);
    P.capx = grid->nx - 2 > 0 ? grid->nx - 2 : 0;
    P.capy = grid->ny - 2 > 0 ? grid->ny - 2 : 0;
    P.capz = grid->nz - 2 > 0 ? grid->nz - 2 : 0;
    std::memcpy(P.g, grid->object_to_grid, sizeof P.g);
    std::memcpy(P.go, grid->grid_origin, sizeof P.go);
    P.views = views;
    P.rows = rows;
    P.cols = cols;
    P.samples = samples;
    P.bundle = panels == nullptr ? 1 : 0;
    configure_tiles(P, 1);
    P.chunk_samples = chunk_samples(samples, grid->nx, grid->ny, grid->nz);
    if (const char *chunk = std::getenv("DPT_AMD_CHUNK")) {
        const int forced = std::atoi(chunk);
        if (forced >= 1 && forced <= 1024) P.chunk_samples = forced;
    }
    P.gauss_chunks = gauss_chunks(grid->nx, grid->ny, grid->nz);
    if (const char *chunks = std::getenv("DPT_AMD_GAUSS_CHUNKS")) {
        const int forced = std::atoi(chunks);
        if (forced >= 1 && forced <= 256) P.gauss_chunks = forced;
    }
    P.slab_chunks = 0;
    P.brick_cap = P.bundle ? kBrickCapDefault : choose_brick_cap(P, panels, &P.slab_chunks);
    if (const char *cap = std::getenv("DPT_AMD_BRICK_CAP")) {
        const int forced = std::atoi(cap);
        if (forced >= 512 && forced <= 8192) P.brick_cap = forced;
    }
    // slab chunks are a form of the FP64 volume adjoint only
    if (const char *slab = std::getenv("DPT_AMD_SLAB_CHUNKS")) {
        if (std::strcmp(slab, "0") == 0) P.slab_chunks = 0;
        if (std::strcmp(slab, "1") == 0) P.slab_chunks = 1;
    }
    p->vjp_vb = P.bundle ? 1 : choose_views_per_block(panels, views, rows, cols);
    if (const char *vb = std::getenv("DPT_AMD_VIEWS_PER_BLOCK")) {
        const int forced = std::atoi(vb);
        if (!P.bundle && (forced == 1 || forced == 2 || forced == 4)) p->vjp_vb = forced;
    }
    p->forward_vb = p->vjp_vb;
    if (const char *vb = std::getenv("DPT_AMD_FORWARD_VIEWS_PER_BLOCK")) {
        const int forced = std::atoi(vb);
        if (!P.bundle && (forced == 1 || forced == 2 || forced == 4)) p->forward_vb = forced;
    }
    p->cells = (needs & DPT_GATHER_CELLS) != 0;
    if (const char *gather = std::getenv("DPT_AMD_GATHER")) {
        if (std::strcmp(gather, "cells") == 0) p->cells = true;
        if (std::strcmp(gather, "linear") == 0) p->cells = false;
    }
    p->cells = p->cells && arithmetic == DPT_ARITH_FP32;
    p->mx = grid->nx > 1 ? grid->nx - 1 : 1;
    p->my = grid->ny > 1 ? grid->ny - 1 : 1;
    p->cell_count = (int64_t)p->mx * p->my * (grid->nz > 1 ? grid->nz - 1 : 1);

    auto cleanup = [&](int code, const std::string &msg) {
        dpt_plan_destroy(p);
        return fail(code, msg);
    };
#define ALLOC(ptr, bytes)                                                                   \
    do {                                                                                    \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess)                              \
            return cleanup(DPT_E_MEMORY, "hipMalloc failed for " #ptr);                     \
    } while (0)

    if (!P.bundle) {
        ALLOC(p->d_panels, sizeof(double) * 12 * views);
        if (hipMemcpy(p->d_panels, panels, sizeof(double) * 12 * views, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(DPT_E_DEVICE, "copy of panels failed");
    } else {
        const size_t bytes = sizeof(double) * 3 * (size_t)pixels;
        ALLOC(p->d_ray_src, bytes);
        ALLOC(p->d_ray_end, bytes);
        if (generated) {
            ALLOC(p->d_panel_meta, sizeof(double) * 18 * views);
            ALLOC(p->d_panel_modes, sizeof(int) * views);
            p->image_cols = image_cols;
            bool ok = hipMemcpy(p->d_panel_meta, panel_meta, sizeof(double) * 18 * views, hipMemcpyHostToDevice) ==
                          hipSuccess &&
                      hipMemcpy(p->d_panel_modes, panel_modes, sizeof(int) * views, hipMemcpyHostToDevice) ==
                          hipSuccess;
            if (ok) {
                hipLaunchKernelGGL(panel_rays_kernel, dim3(grid_blocks(pixels)), dim3(kBlock), 0, 0, p->d_panel_meta,
                                   p->d_panel_modes, (int64_t)cols, image_cols, pixels, p->d_ray_src, p->d_ray_end);
                ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
            }
            if (!ok) return cleanup(DPT_E_DEVICE, "panel ray generation failed");
        } else if (hipMemcpy(p->d_ray_src, ray_sources, bytes, hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy(p->d_ray_end, ray_endpoints, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            return cleanup(DPT_E_DEVICE, "copy of rays failed");
        }
    }
    P.panels = p->d_panels;
    P.ray_src = p->d_ray_src;
    P.ray_end = p->d_ray_end;
    ALLOC(p->d_status, sizeof(int) * views);
    if (hipMemset(p->d_status, 0, sizeof(int) * views) != hipSuccess)
        return cleanup(DPT_E_DEVICE, "status clear failed");
    if (hipHostMalloc((void **)&p->h_status, sizeof(int32_t) * views, hipHostMallocDefault) != hipSuccess)
        return cleanup(DPT_E_MEMORY, "pinned status allocation failed");
    if (needs & DPT_POSE) ALLOC(p->d_partials, sizeof(double) * kMaxComponents * (size_t)P.wt_per_view * views);
    if (needs & DPT_VOLUME) ALLOC(p->d_acc, sizeof(double) * (size_t)voxels);
    if (p->cells) ALLOC(p->d_cells, (storage == DPT_STORE_F32 ? 4 : 2) * 8 * (size_t)p->cell_count);
    ALLOC(p->d_count, sizeof(unsigned long long));
    ALLOC(p->d_alpha, sizeof(float) * sample_table_entries(samples));
    {
        std::vector<float> alpha(sample_table_entries(samples));
        fill_alpha_table(alpha.data(), samples);
        if (hipMemcpy(p->d_alpha, alpha.data(), sizeof(float) * alpha.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(DPT_E_DEVICE, "copy of the sample fractions failed");
    }
    P.alpha = p->d_alpha;
    for (auto &e : p->ev)
        if (hipEventCreate(&e) != hipSuccess) return cleanup(DPT_E_DEVICE, "event creation failed");
#undef ALLOC
    *out = p;
    return DPT_OK;
}

}  // namespace

extern "C" {

int dpt_plan_create(const dpt_grid_desc *grid, int32_t views, int32_t rows, int32_t cols, const double *panels,
                    const double *ray_sources, const double *ray_endpoints, int32_t samples, int32_t arithmetic,
                    int32_t storage, uint32_t needs, int32_t device, dpt_plan **out) {
    return plan_create(grid, views, rows, cols, panels, ray_sources, ray_endpoints, nullptr, nullptr, samples,
                       arithmetic, storage, needs, device, out);
}

int dpt_plan_create_panels(const dpt_grid_desc *grid, int32_t views, int32_t rows, int32_t cols,
                           const double *panel_records, const int32_t *panel_modes, int32_t samples,
                           int32_t arithmetic, int32_t storage, uint32_t needs, int32_t device, dpt_plan **out) {
    if (panel_records == nullptr) return fail(DPT_E_INVALID, "panel records are required");
    return plan_create(grid, views, rows, cols, nullptr, nullptr, nullptr, panel_records, panel_modes, samples,
                       arithmetic, storage, needs, device, out);
}

int dpt_forward(dpt_plan *p, const void *field, const double *poses, void *out, void *stream) {
    if (p == nullptr || field == nullptr || poses == nullptr || out == nullptr)
        return fail(DPT_E_INVALID, "plan, field, poses and out are required");
    DeviceGuard guard(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    Params P = launch_params(p, poses);
    configure_tiles(P, p->forward_vb);
    switch (p->storage) {
    case DPT_STORE_F32:
        hipLaunchKernelGGL(check_field_kernel<float>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0, s,
                           static_cast<const float *>(field), p->voxels, p->d_status);
        break;
    case DPT_STORE_BF16:
        hipLaunchKernelGGL(check_field_kernel<__hip_bfloat16>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0, s,
                           static_cast<const __hip_bfloat16 *>(field), p->voxels, p->d_status);
        break;
    default:
        hipLaunchKernelGGL(check_field_kernel<__half>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0, s,
                           static_cast<const __half *>(field), p->voxels, p->d_status);
    }
    hipLaunchKernelGGL(check_poses_kernel, dim3((P.views + 63) / 64), dim3(64), 0, s, poses, P.views, p->d_status);
    HIP_TRY(hipGetLastError());
    if (p->cells) {
        const int rc = refresh_cells(p, field, s);
        if (rc != DPT_OK) return rc;
    }
    timer_begin(p, 0, s);
    int rc;
    if (p->storage == DPT_STORE_F32)
        rc = launch_forward<float>(p, P, field, out, p->cells, s);
    else if (p->storage == DPT_STORE_BF16)
        rc = launch_forward<__hip_bfloat16>(p, P, field, out, p->cells, s);
    else
        rc = launch_forward<__half>(p, P, field, out, p->cells, s);
    timer_end(p, 0, s);
    return rc;
}

int dpt_vjp(dpt_plan *p, const void *field, const double *poses, const void *cotangent, void *grad_field,
            double *grad_pose, uint32_t flags, void *stream) {
    if (p == nullptr || field == nullptr || poses == nullptr || cotangent == nullptr)
        return fail(DPT_E_INVALID, "plan, field, poses and cotangent are required");
    const bool vol = flags & DPT_VOLUME, pose = flags & DPT_POSE;
    if (!vol && !pose) return fail(DPT_E_INVALID, "request at least one projection gradient");
    if (vol && (grad_field == nullptr || !(p->needs & DPT_VOLUME)))
        return fail(DPT_E_INVALID, "volume gradient requested without a destination or plan support");
    if (pose && (grad_pose == nullptr || !(p->needs & DPT_POSE)))
        return fail(DPT_E_INVALID, "pose gradient requested without a destination or plan support");
    if (flags & ~(uint32_t)(DPT_VOLUME | DPT_POSE | DPT_POSE_TANGENT | DPT_ACCUMULATE | DPT_NO_CHECKS |
                            DPT_VOLUME_FP64))
        return fail(DPT_E_INVALID, "unknown VJP flags");
    const bool vol64 = vol && (flags & DPT_VOLUME_FP64);
    const int components = (flags & DPT_POSE_TANGENT) ? 6 : 12;
    DeviceGuard guard(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    Params P = launch_params(p, poses);
    P.volume_f64 = vol64 ? 1 : 0;
    configure_tiles(P, p->vjp_vb);
    const Params whole = P;
    if (p->arithmetic == DPT_ARITH_FP32 && !p->cell_gauss) configure_segments(P);
    // FP32 storage keeps each ray's pose sums one sequential FP32 sum, as the
    // reference: a segmented launch takes the pose unsegmented through the
    // pipelined loop, beside a segmented volume-only launch
    const bool pipe = pose && P.segments > 1 && p->storage == DPT_STORE_F32;
    const int pose_segments = pipe ? 1 : P.segments;
    if (pose && pose_segments > p->partial_segments) {
        // grown once per plan, for launches small enough to take segments
        HIP_TRY(hipStreamSynchronize(s));
        (void)hipFree(p->d_partials);
        p->d_partials = nullptr;
        p->partial_segments = 0;
        if (hipMalloc((void **)&p->d_partials,
                      sizeof(double) * kMaxComponents * (size_t)P.wt_per_view * P.views * pose_segments) != hipSuccess)
            return fail(DPT_E_MEMORY, "pose partials allocation failed");
        p->partial_segments = pose_segments;
    }
    if (!(flags & DPT_NO_CHECKS)) {
        switch (p->storage) {
        case DPT_STORE_F32:
            hipLaunchKernelGGL(check_field_kernel<float>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0, s,
                               static_cast<const float *>(field), p->voxels, p->d_status);
            break;
        case DPT_STORE_BF16:
            hipLaunchKernelGGL(check_field_kernel<__hip_bfloat16>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0,
                               s, static_cast<const __hip_bfloat16 *>(field), p->voxels, p->d_status);
            break;
        default:
            hipLaunchKernelGGL(check_field_kernel<__half>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0, s,
                               static_cast<const __half *>(field), p->voxels, p->d_status);
        }
        hipLaunchKernelGGL(check_poses_kernel, dim3((P.views + 63) / 64), dim3(64), 0, s, poses, P.views,
                           p->d_status);
        if (p->arithmetic == DPT_ARITH_FP32)
            hipLaunchKernelGGL(check_finite_kernel<float>, dim3(grid_blocks(p->pixels)), dim3(kBlock), 0, s,
                               static_cast<const float *>(cotangent), p->pixels, p->d_status);
        else
            hipLaunchKernelGGL(check_finite_kernel<double>, dim3(grid_blocks(p->pixels)), dim3(kBlock), 0, s,
                               static_cast<const double *>(cotangent), p->pixels, p->d_status);
    }
    if (vol)
        hipLaunchKernelGGL(fill_f64_kernel, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0, s, p->d_acc, p->voxels);
    HIP_TRY(hipGetLastError());
    // the cell copy is rebuilt at every call: the field of this VJP may differ
    // from the last forward's
    const bool cells = p->cells && pose;
    if (cells) {
        const int rc = refresh_cells(p, field, s);
        if (rc != DPT_OK) return rc;
    }
    timer_begin(p, 1, s);
    int rc = DPT_OK;
    if (pipe) {
        if (vol) rc = launch_vjp<float>(p, P, field, cotangent, true, false, components, false, s);
        if (rc == DPT_OK) rc = launch_vjp<float>(p, whole, field, cotangent, false, true, components, cells, s, true);
    } else if (p->storage == DPT_STORE_F32)
        rc = launch_vjp<float>(p, P, field, cotangent, vol, pose, components, cells, s);
    else if (p->storage == DPT_STORE_BF16)
        rc = launch_vjp<__hip_bfloat16>(p, P, field, cotangent, vol, pose, components, cells, s);
    else
        rc = launch_vjp<__half>(p, P, field, cotangent, vol, pose, components, cells, s);
    timer_end(p, 1, s);
    if (rc != DPT_OK) return rc;
    const int accumulate = (flags & DPT_ACCUMULATE) ? 1 : 0;
    if (pose) {
        if (components == 6)
            hipLaunchKernelGGL(pose_reduce_kernel<6>, dim3(P.views), dim3(kBlock), 0, s, p->d_partials,
                               pose_segments * P.wt_per_view, grad_pose, accumulate, p->d_status);
        else
            hipLaunchKernelGGL(pose_reduce_kernel<12>, dim3(P.views), dim3(kBlock), 0, s, p->d_partials,
                               pose_segments * P.wt_per_view, grad_pose, accumulate, p->d_status);
    }
    if (vol64)
        hipLaunchKernelGGL(finish_volume_kernel<double>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0, s,
                           p->d_acc, static_cast<double *>(grad_field), p->voxels, accumulate, p->d_status);
    else if (vol)
        hipLaunchKernelGGL(finish_volume_kernel<float>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0, s,
                           p->d_acc, static_cast<float *>(grad_field), p->voxels, accumulate, p->d_status);
    HIP_TRY(hipGetLastError());
    return DPT_OK;
}

}  // extern "C"

namespace {

// Per-ray VJP outputs (6 FP64 values per ray into the caller's buffer):
// mode kRayCoordinates (world source / endpoint derivatives) or kRayTangent
// (6-DoF local-right pose tangent).
int per_ray_vjp(dpt_plan *p, const void *field, const double *poses, const void *cotangent, double *out_rays,
                int mode, uint32_t flags, void *stream) {
    DeviceGuard guard(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    Params P = launch_params(p, poses);
    configure_tiles(P, p->vjp_vb);
    if (!(flags & DPT_NO_CHECKS)) {
        switch (p->storage) {
        case DPT_STORE_F32:
            hipLaunchKernelGGL(check_field_kernel<float>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0, s,
                               static_cast<const float *>(field), p->voxels, p->d_status);
            break;
        case DPT_STORE_BF16:
            hipLaunchKernelGGL(check_field_kernel<__hip_bfloat16>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0,
                               s, static_cast<const __hip_bfloat16 *>(field), p->voxels, p->d_status);
            break;
        default:
            hipLaunchKernelGGL(check_field_kernel<__half>, dim3(grid_blocks(p->voxels)), dim3(kBlock), 0, s,
                               static_cast<const __half *>(field), p->voxels, p->d_status);
        }
        hipLaunchKernelGGL(check_poses_kernel, dim3((P.views + 63) / 64), dim3(64), 0, s, poses, P.views,
                           p->d_status);
        if (p->arithmetic == DPT_ARITH_FP32)
            hipLaunchKernelGGL(check_finite_kernel<float>, dim3(grid_blocks(p->pixels)), dim3(kBlock), 0, s,
                               static_cast<const float *>(cotangent), p->pixels, p->d_status);
        else
            hipLaunchKernelGGL(check_finite_kernel<double>, dim3(grid_blocks(p->pixels)), dim3(kBlock), 0, s,
                               static_cast<const double *>(cotangent), p->pixels, p->d_status);
    }
    HIP_TRY(hipGetLastError());
    if (p->cells) {
        const int rc = refresh_cells(p, field, s);
        if (rc != DPT_OK) return rc;
    }
    p->ray_out = out_rays;
    int rc;
    if (p->storage == DPT_STORE_F32)
        rc = launch_vjp<float>(p, P, field, cotangent, false, true, mode, p->cells, s);
    else if (p->storage == DPT_STORE_BF16)
        rc = launch_vjp<__hip_bfloat16>(p, P, field, cotangent, false, true, mode, p->cells, s);
    else
        rc = launch_vjp<__half>(p, P, field, cotangent, false, true, mode, p->cells, s);
    p->ray_out = nullptr;
    return rc;
}

}  // namespace

extern "C" {

int dpt_ray_vjp(dpt_plan *p, const void *field, const double *poses, const void *cotangent, double *out_rays,
                uint32_t flags, void *stream) {
    if (p == nullptr || field == nullptr || poses == nullptr || cotangent == nullptr || out_rays == nullptr)
        return fail(DPT_E_INVALID, "plan, field, poses, cotangent and out_rays are required");
    if (flags & ~(uint32_t)DPT_NO_CHECKS) return fail(DPT_E_INVALID, "ray-coordinate VJP flags are DPT_NO_CHECKS");
    return per_ray_vjp(p, field, poses, cotangent, out_rays, kRayCoordinates, flags, stream);
}

int dpt_pose_sensitivities(dpt_plan *p, const void *field, const double *poses, const void *cotangent,
                           double *out_jacobian, uint32_t flags, void *stream) {
    if (p == nullptr || field == nullptr || poses == nullptr || cotangent == nullptr || out_jacobian == nullptr)
        return fail(DPT_E_INVALID, "plan, field, poses, cotangent and out_jacobian are required");
    if (flags & ~(uint32_t)DPT_NO_CHECKS) return fail(DPT_E_INVALID, "pose sensitivity flags are DPT_NO_CHECKS");
    return per_ray_vjp(p, field, poses, cotangent, out_jacobian, kRayTangent, flags, stream);
}

}  // extern "C"

namespace {

// Per-view sum of ray-coordinate derivatives contracted with the view's
// acquisition Jacobian: out[v, k] = sum_p sum_j rays[v, p, j] * J[v, k, j]
// (kernels/batched_projection.py:15-89, fixed-order FP64 tree per view).
__global__ __launch_bounds__(kBlock) void rays_reduce_kernel(const double *__restrict__ rays,
                                                             const double *__restrict__ jacobians, int64_t pixels,
                                                             int components, double *out, int accumulate,
                                                             int *status) {
    const int view = blockIdx.x;
    __shared__ double red[kBlock];
    const double *base = rays + (size_t)view * pixels * 6;
    for (int k = 0; k < components; ++k) {
        const double *jac = jacobians + ((size_t)view * components + k) * 6;
        double v = 0.0;
        for (int64_t i = threadIdx.x; i < pixels; i += kBlock) {
            const double *g = base + 6 * (size_t)i;
            v += g[0] * jac[0] + g[1] * jac[1] + g[2] * jac[2] + g[3] * jac[3] + g[4] * jac[4] + g[5] * jac[5];
        }
        red[threadIdx.x] = v;
        __syncthreads();
        for (int stride = kBlock / 2; stride > 0; stride >>= 1) {
            if (threadIdx.x < stride) red[threadIdx.x] += red[threadIdx.x + stride];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double result = red[0];
            if (accumulate) result += out[(size_t)view * components + k];
            out[(size_t)view * components + k] = result;
            if (!isfinite(result)) atomicOr(&status[view], DPT_STATUS_RANGE);
        }
        __syncthreads();
    }
}

}  // namespace

extern "C" {

int dpt_reduce_rays(dpt_plan *p, const double *rays, const double *jacobians, int32_t components, double *out,
                    uint32_t flags, void *stream) {
    if (p == nullptr || rays == nullptr || jacobians == nullptr || out == nullptr)
        return fail(DPT_E_INVALID, "plan, rays, jacobians and out are required");
    if (components < 1 || components > 16) return fail(DPT_E_INVALID, "components must be 1..16");
    if (flags & ~(uint32_t)DPT_ACCUMULATE) return fail(DPT_E_INVALID, "reduction flags are DPT_ACCUMULATE");
    DeviceGuard guard(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t per_view = p->pixels / p->base.views;
    hipLaunchKernelGGL(rays_reduce_kernel, dim3(p->base.views), dim3(kBlock), 0, s, rays, jacobians, per_view,
                       components, out, (flags & DPT_ACCUMULATE) ? 1 : 0, p->d_status);
    HIP_TRY(hipGetLastError());
    return DPT_OK;
}

int dpt_adjoint_gather(dpt_plan *p, const double *poses, const double *cotangent, void *out, int32_t out_precision,
                       uint32_t flags, void *stream) {
    if (p == nullptr || poses == nullptr || cotangent == nullptr || out == nullptr)
        return fail(DPT_E_INVALID, "plan, poses, cotangent and out are required");
    if (p->arithmetic != DPT_ARITH_FP64 || p->cell_gauss)
        return fail(DPT_E_INVALID, "deterministic gather requires strict float64 midpoint sampling");
    if (out_precision != DPT_ARITH_FP32 && out_precision != DPT_ARITH_FP64)
        return fail(DPT_E_INVALID, "out_precision must be DPT_ARITH_FP32 or DPT_ARITH_FP64");
    if (flags & ~(uint32_t)(DPT_ACCUMULATE | DPT_NO_CHECKS))
        return fail(DPT_E_INVALID, "gather flags are DPT_ACCUMULATE and DPT_NO_CHECKS");
    DeviceGuard guard(p->device);
    if (p->d_gather_rays == nullptr) {
        if (hipMalloc((void **)&p->d_gather_rays, sizeof(GatherRay) * (size_t)p->pixels) != hipSuccess)
            return fail(DPT_E_MEMORY, "hipMalloc failed for the gather ray table");
        if (hipMalloc((void **)&p->d_gather_views, sizeof(GatherView) * (size_t)p->base.views) != hipSuccess ||
            hipMalloc((void **)&p->d_guard, sizeof(int)) != hipSuccess) {
            (void)hipFree(p->d_gather_rays);
            (void)hipFree(p->d_gather_views);
            p->d_gather_rays = nullptr;
            p->d_gather_views = nullptr;
            return fail(DPT_E_MEMORY, "hipMalloc failed for the gather view table");
        }
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Params P = launch_params(p, poses);
    timer_begin(p, 2, s);
    HIP_TRY(launch_gather(P, cotangent, p->d_gather_rays, p->d_gather_views, p->d_guard, out,
                          out_precision == DPT_ARITH_FP64, (flags & DPT_ACCUMULATE) != 0,
                          (flags & DPT_NO_CHECKS) == 0, p->d_panel_meta, p->d_panel_modes, p->image_cols, s));
    timer_end(p, 2, s);
    return DPT_OK;
}

}  // extern "C"

// ------------------------------------------------ allocation-free gather --
// Workspace of dpt_adjoint_gather_noalloc: the gather's ray table, view
// table and guard word in one file-scope device array of fixed size.  Every
// GPU the library runs on holds its own copy until the process ends; in host
// code `scratch` names a host-side stand-in, so kernels and copies only ever
// get the per-device address that hipGetSymbolAddress returns.
namespace {

constexpr size_t kScratchBytes = (size_t)DPT_GATHER_NOALLOC_BYTES;
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_lock;
unsigned char *g_scratch[kScratchDevices] = {};

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// The device address of this device's scratch (looked up once per device,
// with that device current).
int scratch_address(int device, unsigned char **out) {
    if (device < 0 || device >= kScratchDevices) return fail(DPT_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (g_scratch[device] == nullptr) {
        DeviceGuard guard(device);
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (ptr == nullptr) return fail(DPT_E_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch[device] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch[device];
    return DPT_OK;
}

}  // namespace

extern "C" {

int dpt_adjoint_gather_noalloc(dpt_plan *p, const double *poses, const double *cotangent, void *out,
                               int32_t out_precision, uint32_t flags, void *stream) {
    if (p == nullptr || poses == nullptr || cotangent == nullptr || out == nullptr)
        return fail(DPT_E_INVALID, "plan, poses, cotangent and out are required");
    if (p->arithmetic != DPT_ARITH_FP64 || p->cell_gauss)
        return fail(DPT_E_INVALID, "deterministic gather requires strict float64 midpoint sampling");
    if (out_precision != DPT_ARITH_FP32 && out_precision != DPT_ARITH_FP64)
        return fail(DPT_E_INVALID, "out_precision must be DPT_ARITH_FP32 or DPT_ARITH_FP64");
    if (flags & ~(uint32_t)(DPT_ACCUMULATE | DPT_NO_CHECKS))
        return fail(DPT_E_INVALID, "gather flags are DPT_ACCUMULATE and DPT_NO_CHECKS");
    // [ray table | view table | guard], each 256-byte aligned
    const size_t ray_bytes = align256(sizeof(GatherRay) * (size_t)p->pixels);
    const size_t view_bytes = align256(sizeof(GatherView) * (size_t)p->base.views);
    const size_t need = ray_bytes + view_bytes + 256;
    if (need > kScratchBytes)
        return fail(DPT_E_INVALID, "the allocation-free gather needs " + std::to_string(need) +
                                       " bytes of workspace for " + std::to_string(p->pixels) + " rays and " +
                                       std::to_string(p->base.views) + " views; its fixed workspace holds " +
                                       std::to_string(kScratchBytes) + " (use dpt_adjoint_gather)");
    unsigned char *base = nullptr;
    const int rc = scratch_address(p->device, &base);
    if (rc != DPT_OK) return rc;
    DeviceGuard guard(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    GatherRay *rays = reinterpret_cast<GatherRay *>(base);
    GatherView *views = reinterpret_cast<GatherView *>(base + ray_bytes);
    int *guard_word = reinterpret_cast<int *>(base + ray_bytes + view_bytes);
    // launch_gather clears the guard word with a kernel and rewrites both
    // tables before the voxel pass reads them
    HIP_TRY(launch_gather(launch_params(p, poses), cotangent, rays, views, guard_word, out,
                          out_precision == DPT_ARITH_FP64, (flags & DPT_ACCUMULATE) != 0,
                          (flags & DPT_NO_CHECKS) == 0, p->d_panel_meta, p->d_panel_modes, p->image_cols, s));
    return DPT_OK;
}

int dpt_plan_layout(const dpt_plan *p, int32_t *out) {
    if (p == nullptr || out == nullptr) return fail(DPT_E_INVALID, "plan and out are required");
    out[0] = p->base.brick_cap;
    out[1] = p->base.slab_chunks;
    out[2] = p->base.chunk_samples;
    out[3] = p->vjp_vb;
    out[4] = p->base.gauss_chunks;
    return DPT_OK;
}

int dpt_status(dpt_plan *p, int32_t *host_codes, int32_t clear, void *stream) {
    if (p == nullptr || host_codes == nullptr) return fail(DPT_E_INVALID, "plan and host_codes are required");
    DeviceGuard guard(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int views = p->base.views;
    HIP_TRY(hipMemcpyAsync(p->h_status, p->d_status, sizeof(int32_t) * views, hipMemcpyDeviceToHost, s));
    if (clear) hipLaunchKernelGGL(fill_i32_kernel, dim3((views + 255) / 256), dim3(256), 0, s, p->d_status, views);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    std::memcpy(host_codes, p->h_status, sizeof(int32_t) * views);
    return DPT_OK;
}

int dpt_convert_field(const float *src, void *dst, int64_t count, int32_t storage, int32_t *status, void *stream) {
    if (src == nullptr || dst == nullptr || count < 0) return fail(DPT_E_INVALID, "src, dst and count are required");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (storage == DPT_STORE_BF16)
        hipLaunchKernelGGL(convert_kernel<__hip_bfloat16>, dim3(grid_blocks(count)), dim3(kBlock), 0, s, src,
                           static_cast<__hip_bfloat16 *>(dst), count, status);
    else if (storage == DPT_STORE_F16)
        hipLaunchKernelGGL(convert_kernel<__half>, dim3(grid_blocks(count)), dim3(kBlock), 0, s, src,
                           static_cast<__half *>(dst), count, status);
    else
        return fail(DPT_E_INVALID, "storage must be DPT_STORE_BF16 or DPT_STORE_F16");
    HIP_TRY(hipGetLastError());
    return DPT_OK;
}

int64_t dpt_cell_count(int32_t nx, int32_t ny, int32_t nz) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(DPT_E_INVALID, "grid counts must be positive");
    return (int64_t)(nx > 1 ? nx - 1 : 1) * (ny > 1 ? ny - 1 : 1) * (nz > 1 ? nz - 1 : 1);
}

int dpt_convert_cells(const void *src, void *dst, int32_t nx, int32_t ny, int32_t nz, int32_t storage,
                      void *stream) {
    if (src == nullptr || dst == nullptr) return fail(DPT_E_INVALID, "src and dst are required");
    if (nx < 1 || ny < 1 || nz < 1 || (int64_t)nx * ny * nz > INT32_MAX)
        return fail(DPT_E_INVALID, "grid counts must be positive and fit signed 32-bit indexing");
    if (storage != DPT_STORE_F32 && storage != DPT_STORE_BF16 && storage != DPT_STORE_F16)
        return fail(DPT_E_INVALID, "storage must be DPT_STORE_F32, DPT_STORE_BF16 or DPT_STORE_F16");
    return launch_cells(src, dst, nx, ny, nz, storage, static_cast<hipStream_t>(stream));
}

int dpt_set_timing(dpt_plan *p, int32_t enabled) {
    if (p == nullptr) return fail(DPT_E_INVALID, "plan is required");
    p->timing = enabled != 0;
    p->timed[0] = p->timed[1] = p->timed[2] = false;
    return DPT_OK;
}

double dpt_kernel_ms(dpt_plan *p, int32_t kind) {
    if (p == nullptr || kind < 0 || kind > 2 || !p->timed[kind]) return -1.0;
    DeviceGuard guard(p->device);
    if (hipEventSynchronize(p->ev[2 * kind + 1]) != hipSuccess) return -1.0;
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, p->ev[2 * kind], p->ev[2 * kind + 1]) != hipSuccess) return -1.0;
    return (double)ms;
}

int64_t dpt_count_hit_rays(dpt_plan *p, const double *poses, void *stream) {
    if (p == nullptr || poses == nullptr) {
        fail(DPT_E_INVALID, "plan and poses are required");
        return DPT_E_INVALID;
    }
    DeviceGuard guard(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Params P = launch_params(p, poses);
    unsigned long long zero = 0, host = 0;
    if (hipMemcpyAsync(p->d_count, &zero, sizeof zero, hipMemcpyHostToDevice, s) != hipSuccess) return DPT_E_DEVICE;
    hipLaunchKernelGGL(count_hits_kernel, dim3(tile_grid(P)), dim3(kBlock), 0, s, P, p->d_count);
    if (hipMemcpyAsync(&host, p->d_count, sizeof host, hipMemcpyDeviceToHost, s) != hipSuccess) return DPT_E_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return DPT_E_DEVICE;
    return (int64_t)host;
}

}  // extern "C"
