/*
 * CPU restatement of the reference cone-beam projector (TEST INFRASTRUCTURE).
 *
 * This file is the oracle: only tests, the smoke check and the CPU-baseline
 * leg of bench.py may load it, and only as the checker.  The product path
 * (dpt) never calls into it.
 *
 * It restates, operation for operation, two reference kernel families:
 *
 *   FP32 sampling  kernels/projection_fp32.py  (forward :334-374, VJP :377-617)
 *   FP64 sampling  kernels/projection.py       (forward :802-846, VJP :873-1237)
 *
 * with the FP64 ray setup of kernels/projection.py:79-121 and the pose algebra
 * of kernels/geometry.py:25-52, and three more families:
 *
 *   deterministic gather  kernels/projection_adjoint.py:73-170 (serial scatter
 *                         in (view, pixel, sample, corner) order: ref_gather_f64)
 *   exact cell-Gauss      kernels/projection.py:522-746 (ref_forward_gauss_f64,
 *                         ref_vjp_gauss_f64)
 *   ray coordinates       per-ray world source / endpoint cotangents (output
 *                         mode 7 of the VJPs), kernels/projection_fp32.py:570-583
 *                         and kernels/projection.py:1184-1200
 *
 * Build with -ffp-contract=off: every fused
 * multiply-add below is an explicit fmaf/fma, so unmarked products never fuse.
 * Where the reference leaves an expression unmarked but its code generator
 * (Warp, fuse_fp=True) contracts it, this file writes the contracted form
 * explicitly (noted inline); the product kernels use the same explicit forms.
 *
 * Volume gradients are accumulated in FP64 from the FP32-rounded increments
 * the reference scatters (projection_fp32.py:240-252), i.e. the exact sum
 * the reference's FP32 atomics approximate.  Pose gradients are produced per
 * ray in FP64 and summed per view with compensated (Neumaier) summation.
 *
 * Parity pins (see tests/): closed forms from the reference's own tests, an
 * independent explicit-matrix restatement, adjoint identities, central finite
 * differences and analytic ellipsoid chords.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BIT_POSE 1
#define BIT_RANGE 2
#define BIT_DOMAIN 8

typedef struct {
    int32_t nx, ny, nz;          /* grid sample counts, x fastest */
    int32_t views, rows, cols;   /* detector: rows x cols pixels per view */
    int32_t samples;             /* midpoint samples per clipped ray */
    int32_t bundle;              /* 0: panel geometry, 1: explicit rays */
    double g[9];                 /* object_to_grid, row-major: diag(1/h) O^T */
    double grid_origin[3];       /* first sample centre, object mm */
    const double *panels;        /* views x 12: source, origin, col step, row step */
    const double *ray_sources;   /* views x rows*cols x 3 (bundle) */
    const double *ray_endpoints; /* views x rows*cols x 3 (bundle) */
} ref_plan;

typedef struct {
    double origin[3], direction[3];
    double object_source[3], object_direction[3];
    double world_source_relative[3], world_direction[3];
    double length;
} ref_ray;

static void status_or(int32_t *status, int view, int32_t bits) {
#pragma omp atomic
    status[view] |= bits;
}

static void add_f64(double *target, double value) {
#pragma omp atomic
    *target += value;
}

/* kernels/numerics.py:11-56 -- finite factors multiplied through a mantissa /
 * exponent split so an intermediate cannot overflow or underflow early. */
static double product_n(const double *f, int n) {
    double ordinary = 1.0;
    int invalid = 0, zero = 0;
    for (int i = 0; i < n; ++i) {
        ordinary = (i == 0) ? f[0] : ordinary * f[i];
        if (!isfinite(f[i])) invalid = 1;
        if (f[i] == 0.0) zero = 1;
    }
    if (invalid) return ordinary;
    if (zero) return 0.0;
    double mantissa = 1.0;
    int exponent = 0;
    for (int i = 0; i < n; ++i) {
        int e;
        double m = frexp(f[i], &e);
        mantissa = (i == 0) ? m : mantissa * m;
        exponent += e;
    }
    return ldexp(mantissa, exponent);
}

static double product3(double a, double b, double c) {
    double f[3] = {a, b, c};
    return product_n(f, 3);
}

static double product4(double a, double b, double c, double d) {
    double f[4] = {a, b, c, d};
    return product_n(f, 4);
}

/* kernels/projection.py:79-102 (ray_for_pixel) */
static void ray_setup(const ref_plan *p, const double *pose, int view, int pixel, ref_ray *r) {
    double src[3], end[3];
    if (!p->bundle) {
        const double *pn = p->panels + 12 * (size_t)view;
        int row = pixel / p->cols, col = pixel % p->cols;
        for (int i = 0; i < 3; ++i) {
            src[i] = pn[i];
            end[i] = pn[3 + i] + (double)col * pn[6 + i] + (double)row * pn[9 + i];
        }
    } else {
        size_t ray = (size_t)view * p->rows * p->cols + pixel;
        for (int i = 0; i < 3; ++i) {
            src[i] = p->ray_sources[3 * ray + i];
            end[i] = p->ray_endpoints[3 * ray + i];
        }
    }
    for (int i = 0; i < 3; ++i) {
        r->world_source_relative[i] = src[i] - pose[9 + i];
        r->world_direction[i] = end[i] - src[i];
    }
    /* inverse = R^T, R row-major in pose[0..8] */
    for (int i = 0; i < 3; ++i) {
        r->object_source[i] = pose[i] * r->world_source_relative[0] +
                              pose[3 + i] * r->world_source_relative[1] +
                              pose[6 + i] * r->world_source_relative[2];
        r->object_direction[i] = pose[i] * r->world_direction[0] +
                                 pose[3 + i] * r->world_direction[1] +
                                 pose[6 + i] * r->world_direction[2];
    }
    double rel[3];
    for (int i = 0; i < 3; ++i) rel[i] = r->object_source[i] - p->grid_origin[i];
    for (int i = 0; i < 3; ++i) {
        r->origin[i] = p->g[3 * i] * rel[0] + p->g[3 * i + 1] * rel[1] + p->g[3 * i + 2] * rel[2];
        r->direction[i] = p->g[3 * i] * r->object_direction[0] +
                          p->g[3 * i + 1] * r->object_direction[1] +
                          p->g[3 * i + 2] * r->object_direction[2];
    }
    r->length = sqrt(r->world_direction[0] * r->world_direction[0] +
                     r->world_direction[1] * r->world_direction[1] +
                     r->world_direction[2] * r->world_direction[2]);
}

static int finite3(const double *v) { return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]); }
static int finite3f(const float *v) { return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]); }

/* kernels/projection.py:110-121 (valid_ray) */
static int valid_ray(const ref_ray *r) {
    return finite3(r->origin) && finite3(r->direction) && finite3(r->object_source) &&
           finite3(r->object_direction) && finite3(r->world_source_relative) &&
           finite3(r->world_direction) && isfinite(r->length) && r->length > 0.0;
}

/* ---------------------------------------------------------------- FP32 --- */

typedef struct {
    float lower, upper;
    float lo_o[3], lo_d[3], up_o[3], up_d[3];
} interval32;

/* kernels/projection_fp32.py:77-109 */
static interval32 finite_interval32(const float *o, const float *d, const int *shape) {
    interval32 iv;
    memset(&iv, 0, sizeof iv);
    iv.lower = 0.0f;
    iv.upper = 1.0f;
    for (int a = 0; a < 3; ++a) {
        float low = -0.5f, high = (float)shape[a] - 0.5f, v = d[a];
        if (v == 0.0f) {
            if (o[a] < low || o[a] > high) iv.upper = -1.0f;
        } else {
            float first = (low - o[a]) / v, last = (high - o[a]) / v;
            if (first > last) {
                float t = first;
                first = last;
                last = t;
            }
            if (first > iv.lower) {
                iv.lower = first;
                memset(iv.lo_o, 0, sizeof iv.lo_o);
                memset(iv.lo_d, 0, sizeof iv.lo_d);
                iv.lo_o[a] = -1.0f / v;
                iv.lo_d[a] = -first / v;
            }
            if (last < iv.upper) {
                iv.upper = last;
                memset(iv.up_o, 0, sizeof iv.up_o);
                memset(iv.up_d, 0, sizeof iv.up_d);
                iv.up_o[a] = -1.0f / v;
                iv.up_d[a] = -last / v;
            }
        }
    }
    return iv;
}

/* kernels/projection_fp32.py:64-74 */
static int inside32(const float *p, const int *shape) {
    return finite3f(p) && p[0] >= -0.5f && p[1] >= -0.5f && p[2] >= -0.5f &&
           p[0] <= (float)shape[0] - 0.5f && p[1] <= (float)shape[1] - 0.5f &&
           p[2] <= (float)shape[2] - 0.5f;
}

typedef struct {
    int lower[3];
    float frac[3], slope[3];
} coords32;

static float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

/* kernels/projection_fp32.py:130-140 */
static coords32 coordinates32(const float *p, const int *shape) {
    coords32 c;
    for (int a = 0; a < 3; ++a) {
        float high = (float)(shape[a] - 1);
        float value = clampf(p[a], 0.0f, high);
        int lo = (int)floorf(value);
        int cap = shape[a] - 2 > 0 ? shape[a] - 2 : 0;
        c.lower[a] = lo < cap ? lo : cap;
        c.frac[a] = value - (float)c.lower[a];
        c.slope[a] = (p[a] >= 0.0f && p[a] < high) ? 1.0f : 0.0f;
    }
    return c;
}

/* kernels/projection.py:217-222 */
static size_t corner_index(const int *lower, const int *shape, int corner) {
    int x = lower[0] + (corner & 1), y = lower[1] + ((corner >> 1) & 1), z = lower[2] + ((corner >> 2) & 1);
    if (x > shape[0] - 1) x = shape[0] - 1;
    if (y > shape[1] - 1) y = shape[1] - 1;
    if (z > shape[2] - 1) z = shape[2] - 1;
    return ((size_t)z * shape[1] + y) * shape[0] + x;
}

/* kernels/projection_fp32.py:143-154: weight = (f0*f1)*f2 */
static float corner_weight32(const float *frac, int corner) {
    float f[3];
    for (int a = 0; a < 3; ++a) f[a] = ((corner >> a) & 1) ? frac[a] : 1.0f - frac[a];
    return (f[0] * f[1]) * f[2];
}

/* kernels/projection_fp32.py:157-177 (value only) */
static float sample32(const float *field, const float *p, const int *shape) {
    if (!inside32(p, shape)) return 0.0f;
    coords32 c = coordinates32(p, shape);
    float value = 0.0f;
    for (int k = 0; k < 8; ++k)
        value = fmaf(corner_weight32(c.frac, k), field[corner_index(c.lower, shape, k)], value);
    return value;
}

/* kernels/projection_fp32.py:180-209 */
static float sample_pose32(const float *field, const float *p, const int *shape, float *grad) {
    grad[0] = grad[1] = grad[2] = 0.0f;
    if (!inside32(p, shape)) return 0.0f;
    coords32 c = coordinates32(p, shape);
    float v[8], value = 0.0f;
    for (int k = 0; k < 8; ++k) {
        v[k] = field[corner_index(c.lower, shape, k)];
        value = fmaf(corner_weight32(c.frac, k), v[k], value);
    }
    float fx = c.frac[0], fy = c.frac[1], fz = c.frac[2];
    float ax = 1.0f - fx, ay = 1.0f - fy, az = 1.0f - fz;
    float x0 = fmaf(fy, v[3] - v[2], ay * (v[1] - v[0]));
    float x1 = fmaf(fy, v[7] - v[6], ay * (v[5] - v[4]));
    grad[0] = c.slope[0] * fmaf(fz, x1, az * x0);
    float y0 = fmaf(fx, v[3] - v[1], ax * (v[2] - v[0]));
    float y1 = fmaf(fx, v[7] - v[5], ax * (v[6] - v[4]));
    grad[1] = c.slope[1] * fmaf(fz, y1, az * y0);
    float z0 = fmaf(fx, v[5] - v[1], ax * (v[4] - v[0]));
    float z1 = fmaf(fx, v[7] - v[3], ax * (v[6] - v[2]));
    grad[2] = c.slope[2] * fmaf(fy, z1, ay * z0);
    return value;
}

/* kernels/projection_fp32.py:240-252: RN(factor*weight), |x| < 2^-126 -> signed 0 */
static float scatter_increment(float factor, float weight) {
    volatile float inc = factor * weight;
    float x = inc;
    if (fabsf(x) < 0x1p-126f) return copysignf(0.0f, x);
    return x;
}

static void ray32(const ref_ray *pr, float *o, float *d, float *len) {
    for (int i = 0; i < 3; ++i) {
        o[i] = (float)pr->origin[i];
        d[i] = (float)pr->direction[i];
    }
    *len = (float)pr->length;
}

static int bad_ray32(const ref_ray *pr, const float *o, const float *d, float len) {
    return !valid_ray(pr) || !finite3f(o) || !finite3f(d) || !isfinite(len) || len <= 0.0f;
}

/* kernels/projection_fp32.py:334-374 */
void ref_forward_f32(const ref_plan *p, const float *field, const double *poses, float *out,
                     int32_t *status) {
    const int shape[3] = {p->nx, p->ny, p->nz};
    const int pixels = p->rows * p->cols;
    const long total_rays = (long)p->views * pixels;
#pragma omp parallel for schedule(dynamic, 64)
    for (long ray = 0; ray < total_rays; ++ray) {
        int view = (int)(ray / pixels), pixel = (int)(ray % pixels);
        ref_ray pr;
        ray_setup(p, poses + 12 * (size_t)view, view, pixel, &pr);
        float o[3], d[3], len;
        ray32(&pr, o, d, &len);
        if (bad_ray32(&pr, o, d, len)) {
            status_or(status, view, BIT_RANGE);
            out[ray] = 0.0f;
            continue;
        }
        interval32 iv = finite_interval32(o, d, shape);
        float total = 0.0f;
        if (iv.upper > iv.lower) {
            float step = (iv.upper - iv.lower) / (float)p->samples;
            for (int k = 0; k < p->samples; ++k) {
                float pos = fmaf((float)k + 0.5f, step, iv.lower);
                float pt[3] = {fmaf(pos, d[0], o[0]), fmaf(pos, d[1], o[1]), fmaf(pos, d[2], o[2])};
                if (!finite3f(pt)) status_or(status, view, BIT_RANGE);
                total += sample32(field, pt, shape);
            }
            total *= len * step;
        }
        out[ray] = total;
        if (!isfinite(total)) status_or(status, view, BIT_RANGE);
    }
}

/* Pose epilogue shared by both arithmetics: kernels/projection_fp32.py:562-602
 * and kernels/projection.py:1156-1217, in FP64. */
static void pose_epilogue(const ref_plan *p, const double *pose, const ref_ray *pr, double factor,
                          const double *origin_gradient, const double *direction_gradient,
                          int components, double *out) {
    double sa[3], da[3];
    for (int i = 0; i < 3; ++i) {
        /* G^T * v : (G^T)[i][j] = G[j][i] */
        double s = p->g[i] * origin_gradient[0] + p->g[3 + i] * origin_gradient[1] +
                   p->g[6 + i] * origin_gradient[2];
        double t = p->g[i] * direction_gradient[0] + p->g[3 + i] * direction_gradient[1] +
                   p->g[6 + i] * direction_gradient[2];
        sa[i] = factor * s;
        da[i] = factor * t;
    }
    if (components == 12) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                out[3 * r + c] = pr->world_source_relative[r] * sa[c] + pr->world_direction[r] * da[c];
        for (int a = 0; a < 3; ++a)
            out[9 + a] = -(pose[3 * a] * sa[0] + pose[3 * a + 1] * sa[1] + pose[3 * a + 2] * sa[2]);
    } else {
        /* kernels/geometry.py:25-52: (-sa, sa x os + da x od) */
        const double *os = pr->object_source, *od = pr->object_direction;
        out[0] = -sa[0];
        out[1] = -sa[1];
        out[2] = -sa[2];
        out[3] = (sa[1] * os[2] - sa[2] * os[1]) + (da[1] * od[2] - da[2] * od[1]);
        out[4] = (sa[2] * os[0] - sa[0] * os[2]) + (da[2] * od[0] - da[0] * od[2]);
        out[5] = (sa[0] * os[1] - sa[1] * os[0]) + (da[0] * od[1] - da[1] * od[0]);
    }
}

/* Ray-coordinate mode (components == REF_RAY_COORDINATES): per ray
 * cotangent x dL/d(world source, world endpoint), 6 values
 * (kernels/projection_fp32.py:570-583, kernels/projection.py:1184-1200):
 * L = |e - s| I(s, e - s), so the endpoint gains seed * I * (e - s)/|e - s|;
 * the FP64 path forms that product range-safely. */
#define REF_RAY_COORDINATES 7

static void ray_epilogue(const ref_plan *p, const double *pose, const ref_ray *pr, double factor,
                         const double *og, const double *dg, double seed, double integral, int fp64_form,
                         double *out) {
    double sa[3], da[3], sw[3], dw[3];
    for (int i = 0; i < 3; ++i) {
        const double s = p->g[i] * og[0] + p->g[3 + i] * og[1] + p->g[6 + i] * og[2];
        const double t = p->g[i] * dg[0] + p->g[3 + i] * dg[1] + p->g[6 + i] * dg[2];
        sa[i] = factor * s;
        da[i] = factor * t;
    }
    for (int i = 0; i < 3; ++i) {
        sw[i] = pose[3 * i] * sa[0] + pose[3 * i + 1] * sa[1] + pose[3 * i + 2] * sa[2];
        dw[i] = pose[3 * i] * da[0] + pose[3 * i + 1] * da[1] + pose[3 * i + 2] * da[2];
    }
    for (int a = 0; a < 3; ++a) {
        dw[a] += fp64_form ? product3(seed, integral, pr->world_direction[a] / pr->length)
                           : seed * integral * pr->world_direction[a] / pr->length;
        out[a] = sw[a] - dw[a];
        out[a + 3] = dw[a];
    }
}

static int outputs_of(int components) { return components == REF_RAY_COORDINATES ? 6 : components; }

static void neumaier(double *sum, double *comp, double x) {
    double t = *sum + x;
    if (fabs(*sum) >= fabs(x))
        *comp += (*sum - t) + x;
    else
        *comp += (x - t) + *sum;
    *sum = t;
}

/* Sum per-ray pose gradients of every view (compensated, fixed order). */
static void reduce_pose(const ref_plan *p, const double *per_ray, int components, double *grad_pose) {
    const int pixels = p->rows * p->cols;
    for (int v = 0; v < p->views; ++v) {
        for (int c = 0; c < components; ++c) {
            double s = 0.0, e = 0.0;
            for (int px = 0; px < pixels; ++px)
                neumaier(&s, &e, per_ray[((size_t)v * pixels + px) * components + c]);
            grad_pose[(size_t)v * components + c] = s + e;
        }
    }
}

/*
 * kernels/projection_fp32.py:377-617 (get_vjp).
 * grad_field: FP64 accumulator (voxels) or NULL; grad_pose: views x components
 * or NULL; per_ray: optional views*pixels x components output.
 */
void ref_vjp_f32(const ref_plan *p, const float *field, const double *poses, const float *seeds,
                 double *grad_field, double *grad_pose, double *per_ray_out, int components,
                 int32_t *status) {
    const int shape[3] = {p->nx, p->ny, p->nz};
    const int pixels = p->rows * p->cols;
    const long total_rays = (long)p->views * pixels;
    const int active_pose = grad_pose != NULL || per_ray_out != NULL;
    double *per_ray = per_ray_out;
    if (active_pose && per_ray == NULL)
        per_ray = (double *)calloc((size_t)total_rays * outputs_of(components), sizeof(double));
#pragma omp parallel for schedule(dynamic, 64)
    for (long ray = 0; ray < total_rays; ++ray) {
        int view = (int)(ray / pixels), pixel = (int)(ray % pixels);
        const double *pose = poses + 12 * (size_t)view;
        double gradient[12] = {0};
        ref_ray pr;
        ray_setup(p, pose, view, pixel, &pr);
        float o[3], d[3], len;
        ray32(&pr, o, d, &len);
        if (bad_ray32(&pr, o, d, len)) {
            status_or(status, view, BIT_RANGE);
        } else {
            interval32 iv = finite_interval32(o, d, shape);
            float span = iv.upper - iv.lower;
            float seed = seeds[ray];
            if (!isfinite(seed)) status_or(status, view, BIT_RANGE);
            if (span > 0.0f && seed != 0.0f) {
                float step = span / (float)p->samples;
                float physical_step = len * step;
                float volume_factor = seed * physical_step;
                float mean = 0.0f, dor[3] = {0, 0, 0}, ddir[3] = {0, 0, 0};
                float shift_lower = 0.0f, shift_upper = 0.0f;
                for (int k = 0; k < p->samples; ++k) {
                    float pos = fmaf((float)k + 0.5f, step, iv.lower);
                    float pt[3] = {fmaf(pos, d[0], o[0]), fmaf(pos, d[1], o[1]), fmaf(pos, d[2], o[2])};
                    if (!finite3f(pt)) status_or(status, view, BIT_RANGE);
                    if (active_pose) {
                        float g[3];
                        float value = sample_pose32(field, pt, shape, g);
                        mean += value;
                        for (int a = 0; a < 3; ++a) {
                            dor[a] += g[a];
                            /* contracted by the reference's code generator */
                            ddir[a] = fmaf(pos, g[a], ddir[a]);
                        }
                        float alpha = ((float)k + 0.5f) / (float)p->samples;
                        float along = fmaf(g[2], d[2], fmaf(g[1], d[1], g[0] * d[0]));
                        shift_lower = fmaf(1.0f - alpha, along, shift_lower);
                        shift_upper = fmaf(alpha, along, shift_upper);
                    }
                    if (grad_field != NULL && inside32(pt, shape)) {
                        coords32 c = coordinates32(pt, shape);
                        for (int kk = 0; kk < 8; ++kk) {
                            float inc = scatter_increment(volume_factor, corner_weight32(c.frac, kk));
                            if (!isfinite(inc)) status_or(status, view, BIT_RANGE);
                            add_f64(&grad_field[corner_index(c.lower, shape, kk)], (double)inc);
                        }
                    }
                }
                if (active_pose) {
                    float lower_seed = fmaf(span, shift_lower, -mean);
                    float upper_seed = fmaf(span, shift_upper, mean);
                    double og[3], dg[3];
                    for (int a = 0; a < 3; ++a) {
                        og[a] = (double)fmaf(upper_seed, iv.up_o[a],
                                             fmaf(lower_seed, iv.lo_o[a], span * dor[a]));
                        dg[a] = (double)fmaf(upper_seed, iv.up_d[a],
                                             fmaf(lower_seed, iv.lo_d[a], span * ddir[a]));
                    }
                    float factor = (seed * len) / (float)p->samples;
                    if (components == REF_RAY_COORDINATES)
                        ray_epilogue(p, pose, &pr, (double)factor, og, dg, (double)seed, (double)(mean * step), 0,
                                     gradient);
                    else
                        pose_epilogue(p, pose, &pr, (double)factor, og, dg, components, gradient);
                }
            }
        }
        if (active_pose)
            for (int c = 0; c < outputs_of(components); ++c)
                per_ray[(size_t)ray * outputs_of(components) + c] = gradient[c];
    }
    if (grad_pose != NULL) reduce_pose(p, per_ray, components, grad_pose);
    if (per_ray != per_ray_out) free(per_ray);
}

/* ---------------------------------------------------------------- FP64 --- */

typedef struct {
    double lower, upper;
    double lo_o[3], lo_d[3], up_o[3], up_d[3];
} interval64;

/* kernels/projection.py:147-184 */
static interval64 finite_interval64(const double *o, const double *d, const int *shape) {
    interval64 iv;
    memset(&iv, 0, sizeof iv);
    iv.lower = 0.0;
    iv.upper = 1.0;
    for (int a = 0; a < 3; ++a) {
        double low = -0.5, high = (double)shape[a] - 0.5, v = d[a];
        if (v == 0.0) {
            if (o[a] < low || o[a] > high) iv.upper = -1.0;
        } else {
            double first = (low - o[a]) / v, last = (high - o[a]) / v;
            if (first > last) {
                double t = first;
                first = last;
                last = t;
            }
            if (first > iv.lower) {
                iv.lower = first;
                memset(iv.lo_o, 0, sizeof iv.lo_o);
                memset(iv.lo_d, 0, sizeof iv.lo_d);
                iv.lo_o[a] = -1.0 / v;
                iv.lo_d[a] = -first / v;
            }
            if (last < iv.upper) {
                iv.upper = last;
                memset(iv.up_o, 0, sizeof iv.up_o);
                memset(iv.up_d, 0, sizeof iv.up_d);
                iv.up_o[a] = -1.0 / v;
                iv.up_d[a] = -last / v;
            }
        }
    }
    return iv;
}

static int inside64(const double *p, const int *shape) {
    return finite3(p) && p[0] >= -0.5 && p[1] >= -0.5 && p[2] >= -0.5 &&
           p[0] <= (double)shape[0] - 0.5 && p[1] <= (double)shape[1] - 0.5 &&
           p[2] <= (double)shape[2] - 0.5;
}

static double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

/* kernels/projection.py:199-214 (cell_lower, cell_fraction) */
static void cell64(const double *p, const int *shape, int *lower, double *frac) {
    for (int a = 0; a < 3; ++a) {
        double c = clampd(p[a], 0.0, (double)(shape[a] - 1));
        int lo = (int)floor(c);
        int cap = shape[a] - 2 > 0 ? shape[a] - 2 : 0;
        lower[a] = lo < cap ? lo : cap;
        frac[a] = c - (double)lower[a];
    }
}

/* kernels/projection.py:236-274 (sample_cell_fraction) */
static double sample_cell64(const double *v, const double *p, const int *shape, const double *f,
                            double *grad) {
    double slope[3];
    for (int a = 0; a < 3; ++a) slope[a] = (p[a] >= 0.0 && p[a] < (double)(shape[a] - 1)) ? 1.0 : 0.0;
    double v000 = v[0], v100 = v[1], v010 = v[2], v110 = v[3];
    double v001 = v[4], v101 = v[5], v011 = v[6], v111 = v[7];
    double dx00 = v100 - v000, dx10 = v110 - v010, dx01 = v101 - v001, dx11 = v111 - v011;
    double fx = f[0], fy = f[1], fz = f[2];
    double ax = 1.0 - fx, ay = 1.0 - fy, az = 1.0 - fz;
    double x00 = ax * v000 + fx * v100, x10 = ax * v010 + fx * v110;
    double x01 = ax * v001 + fx * v101, x11 = ax * v011 + fx * v111;
    double xy0 = ay * x00 + fy * x10, xy1 = ay * x01 + fy * x11;
    double value = az * xy0 + fz * xy1;
    if (grad) {
        double dx0 = ay * dx00 + fy * dx10, dx1 = ay * dx01 + fy * dx11;
        double dy0 = ax * (v010 - v000) + fx * (v110 - v100);
        double dy1 = ax * (v011 - v001) + fx * (v111 - v101);
        double dz0 = ax * (v001 - v000) + fx * (v101 - v100);
        double dz1 = ax * (v011 - v010) + fx * (v111 - v110);
        grad[0] = slope[0] * (az * dx0 + fz * dx1);
        grad[1] = slope[1] * (az * dy0 + fz * dy1);
        grad[2] = slope[2] * (ay * dz0 + fy * dz1);
    }
    return value;
}

static void load_cell(const float *field, const int *lower, const int *shape, double *v) {
    for (int k = 0; k < 8; ++k) v[k] = (double)field[corner_index(lower, shape, k)];
}

/* kernels/projection.py:802-846 (midpoint, FP64 sampling) */
void ref_forward_f64(const ref_plan *p, const float *field, const double *poses, double *out,
                     int32_t *status) {
    const int shape[3] = {p->nx, p->ny, p->nz};
    const int pixels = p->rows * p->cols;
    const long total_rays = (long)p->views * pixels;
#pragma omp parallel for schedule(dynamic, 64)
    for (long ray = 0; ray < total_rays; ++ray) {
        int view = (int)(ray / pixels), pixel = (int)(ray % pixels);
        ref_ray r;
        ray_setup(p, poses + 12 * (size_t)view, view, pixel, &r);
        if (!valid_ray(&r)) {
            status_or(status, view, BIT_RANGE);
            out[ray] = 0.0;
            continue;
        }
        interval64 iv = finite_interval64(r.origin, r.direction, shape);
        double integral = 0.0;
        if (iv.upper > iv.lower) {
            double step = (iv.upper - iv.lower) / (double)p->samples;
            for (int k = 0; k < p->samples; ++k) {
                double pos = iv.lower + ((double)k + 0.5) * step;
                double pt[3] = {r.origin[0] + pos * r.direction[0], r.origin[1] + pos * r.direction[1],
                                r.origin[2] + pos * r.direction[2]};
                if (!finite3(pt)) status_or(status, view, BIT_RANGE);
                if (inside64(pt, shape)) {
                    int lower[3];
                    double f[3], v[8];
                    cell64(pt, shape, lower, f);
                    load_cell(field, lower, shape, v);
                    integral += sample_cell64(v, pt, shape, f, NULL);
                }
            }
            integral *= r.length * step;
        }
        out[ray] = integral;
        if (!isfinite(integral)) status_or(status, view, BIT_RANGE);
    }
}

/* kernels/projection.py:873-1237 (midpoint, FP64 sampling, FP32 volume destination) */
void ref_vjp_f64(const ref_plan *p, const float *field, const double *poses, const double *seeds,
                 double *grad_field, double *grad_pose, double *per_ray_out, int components,
                 int32_t *status) {
    const int shape[3] = {p->nx, p->ny, p->nz};
    const int pixels = p->rows * p->cols;
    const long total_rays = (long)p->views * pixels;
    const int active_pose = grad_pose != NULL || per_ray_out != NULL;
    double *per_ray = per_ray_out;
    if (active_pose && per_ray == NULL)
        per_ray = (double *)calloc((size_t)total_rays * outputs_of(components), sizeof(double));
    const int pow2 = p->samples > 0 && (p->samples & (p->samples - 1)) == 0;
#pragma omp parallel for schedule(dynamic, 64)
    for (long ray = 0; ray < total_rays; ++ray) {
        int view = (int)(ray / pixels), pixel = (int)(ray % pixels);
        const double *pose = poses + 12 * (size_t)view;
        double gradient[12] = {0};
        ref_ray r;
        ray_setup(p, pose, view, pixel, &r);
        if (!valid_ray(&r)) {
            status_or(status, view, BIT_RANGE);
        } else {
            interval64 iv = finite_interval64(r.origin, r.direction, shape);
            double span = iv.upper - iv.lower;
            double seed = seeds[ray];
            if (seed == 0.0) span = 0.0;
            if (span > 0.0) {
                double factor = seed * r.length;
                double step = span / (double)p->samples;
                double volume_factor = factor * span / (double)p->samples;
                if (!isfinite(factor))
                    volume_factor = product4(seed, r.length, span, 1.0 / (double)p->samples);
                double inverse_samples = pow2 ? 1.0 / (double)p->samples : 0.0;
                double mean = 0.0, dor[3] = {0, 0, 0}, ddir[3] = {0, 0, 0};
                double shift_lower = 0.0, shift_upper = 0.0;
                for (int k = 0; k < p->samples; ++k) {
                    double alpha = pow2 ? ((double)k + 0.5) * inverse_samples
                                        : ((double)k + 0.5) / (double)p->samples;
                    double pos = iv.lower + ((double)k + 0.5) * step;
                    double pt[3] = {r.origin[0] + pos * r.direction[0],
                                    r.origin[1] + pos * r.direction[1],
                                    r.origin[2] + pos * r.direction[2]};
                    if (!finite3(pt)) status_or(status, view, BIT_RANGE);
                    if (!inside64(pt, shape)) continue;
                    int lower[3];
                    double f[3];
                    cell64(pt, shape, lower, f);
                    if (active_pose) {
                        double v[8], g[3];
                        load_cell(field, lower, shape, v);
                        double value = sample_cell64(v, pt, shape, f, g);
                        mean += value;
                        for (int a = 0; a < 3; ++a) {
                            dor[a] += g[a];
                            ddir[a] += pos * g[a];
                        }
                        double along = g[0] * r.direction[0] + g[1] * r.direction[1] + g[2] * r.direction[2];
                        shift_lower += (1.0 - alpha) * along;
                        shift_upper += alpha * along;
                    }
                    if (grad_field != NULL) {
                        /* kernels/projection.py:396-430: seed*w in axis order, one FP32 rounding */
                        for (int kk = 0; kk < 8; ++kk) {
                            double contribution = volume_factor;
                            for (int a = 0; a < 3; ++a)
                                contribution *= ((kk >> a) & 1) ? f[a] : 1.0 - f[a];
                            float rounded = (float)contribution;
                            if (!isfinite(rounded)) status_or(status, view, BIT_RANGE);
                            add_f64(&grad_field[corner_index(lower, shape, kk)], (double)rounded);
                        }
                    }
                }
                if (active_pose) {
                    double lower_seed = -mean + span * shift_lower;
                    double upper_seed = mean + span * shift_upper;
                    double og[3], dg[3];
                    for (int a = 0; a < 3; ++a) {
                        og[a] = span * dor[a] + lower_seed * iv.lo_o[a] + upper_seed * iv.up_o[a];
                        dg[a] = span * ddir[a] + lower_seed * iv.lo_d[a] + upper_seed * iv.up_d[a];
                    }
                    factor /= (double)p->samples;
                    if (isfinite(factor) && components == REF_RAY_COORDINATES) {
                        ray_epilogue(p, pose, &r, factor, og, dg, seed, mean * step, 1, gradient);
                    } else if (isfinite(factor)) {
                        pose_epilogue(p, pose, &r, factor, og, dg, components, gradient);
                    } else {
                        /* kernels/projection.py:1163-1183: range-safe products */
                        double sg[3], rg[3], sa_unit[3], da_unit[3];
                        for (int i = 0; i < 3; ++i) {
                            sg[i] = p->g[i] * og[0] + p->g[3 + i] * og[1] + p->g[6 + i] * og[2];
                            rg[i] = p->g[i] * dg[0] + p->g[3 + i] * dg[1] + p->g[6 + i] * dg[2];
                        }
                        for (int i = 0; i < 3; ++i) {
                            sa_unit[i] = product4(seed, r.length, 1.0 / (double)p->samples, sg[i]);
                            da_unit[i] = product4(seed, r.length, 1.0 / (double)p->samples, rg[i]);
                        }
                        /* identity G for the epilogue: pass pre-scaled adjoints */
                        ref_plan unit = *p;
                        memset(unit.g, 0, sizeof unit.g);
                        unit.g[0] = unit.g[4] = unit.g[8] = 1.0;
                        if (components == REF_RAY_COORDINATES)
                            ray_epilogue(&unit, pose, &r, 1.0, sa_unit, da_unit, seed, mean * step, 1, gradient);
                        else
                            pose_epilogue(&unit, pose, &r, 1.0, sa_unit, da_unit, components, gradient);
                    }
                }
            }
        }
        if (active_pose)
            for (int c = 0; c < outputs_of(components); ++c)
                per_ray[(size_t)ray * outputs_of(components) + c] = gradient[c];
    }
    if (grad_pose != NULL) reduce_pose(p, per_ray, components, grad_pose);
    if (per_ray != per_ray_out) free(per_ray);
    (void)product3;
}

int ref_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Thread count of later parallel regions (CPU-baseline timing at 1 thread). */
void ref_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/*
 * Deterministic adjoint gather, kernels/projection_adjoint.py:73-170
 * (get_gather, midpoint): out[v] = sum over views, pixels (row-major),
 * samples and corners 0..7, in that order, of
 * scaled_corner_weight(fraction, corner, volume_factor) (kernels/projection.py:
 * 396-404: ((vf * w_x) * w_y) * w_z with no FP32 rounding) for every corner
 * whose clamped index is v; FP64 throughout.
 *
 * The reference computes each voxel's sum by visiting every ray in that
 * order.  A serial scatter over the same (view, pixel, sample, corner) order
 * hands every voxel exactly the same sequence of addends, so it produces the
 * same FP64 sums bit for bit at a cost of rays x samples instead of
 * voxels x rays x samples; tests pin it against a voxel-driven Python
 * restatement.  Single-threaded on purpose: the order is the contract.
 *
 * Rays skipped as in the reference: invalid (status RANGE), zero cotangent,
 * empty clipped interval.  A non-finite sample point, or a non-finite final
 * value, sets RANGE (the latter on status[0]).  With accumulate the previous
 * out[v] is added after the local sum (projection_adjoint.py:164-166).
 */
void ref_gather_f64(const ref_plan *p, const double *poses, const double *cotangent, double *out,
                    int accumulate, int32_t *status) {
    const int shape[3] = {p->nx, p->ny, p->nz};
    const int pixels = p->rows * p->cols;
    const long total_rays = (long)p->views * pixels;
    const size_t voxels = (size_t)p->nx * p->ny * p->nz;
    double *total = (double *)calloc(voxels, sizeof(double));
    for (long ray = 0; ray < total_rays; ++ray) {
        int view = (int)(ray / pixels), pixel = (int)(ray % pixels);
        ref_ray r;
        ray_setup(p, poses + 12 * (size_t)view, view, pixel, &r);
        if (!valid_ray(&r)) {
            status[view] |= BIT_RANGE;
            continue;
        }
        const double seed = cotangent[ray];
        if (seed == 0.0) continue;
        interval64 iv = finite_interval64(r.origin, r.direction, shape);
        const double span = iv.upper - iv.lower;
        if (span <= 0.0) continue;
        const double factor = seed * r.length;
        const double step = span / (double)p->samples;
        double volume_factor = factor * span / (double)p->samples;
        if (!isfinite(factor)) volume_factor = product4(seed, r.length, span, 1.0 / (double)p->samples);
        for (int k = 0; k < p->samples; ++k) {
            const double pos = iv.lower + ((double)k + 0.5) * step;
            const double pt[3] = {r.origin[0] + pos * r.direction[0], r.origin[1] + pos * r.direction[1],
                                  r.origin[2] + pos * r.direction[2]};
            if (!finite3(pt)) status[view] |= BIT_RANGE;
            if (!inside64(pt, shape)) continue;
            int lower[3];
            double f[3];
            cell64(pt, shape, lower, f);
            for (int corner = 0; corner < 8; ++corner) {
                double contribution = volume_factor;
                for (int a = 0; a < 3; ++a) contribution *= ((corner >> a) & 1) ? f[a] : 1.0 - f[a];
                total[corner_index(lower, shape, corner)] += contribution;
            }
        }
    }
    for (size_t v = 0; v < voxels; ++v) {
        double value = total[v];
        if (accumulate) value += out[v];
        out[v] = value;
        if (!isfinite(value)) status[0] |= BIT_RANGE;
    }
    free(total);
}

/* ------------------------------------------------------- cell-Gauss --- */
/*
 * Exact cell-Gauss integration (kernels/projection.py:522-746): the clipped
 * ray is cut at every interpolation plane it crosses (integer coordinates),
 * and each piece is integrated with two-point Gauss nodes, which is exact
 * for the cubic that a trilinear field is along a line inside one cell.
 */

typedef struct {
    int plane[3], step[3];
    double next[3], inv[3];
} cell_walk;

/* kernels/projection.py:629-640 (crossing_quotient): the cached reciprocal
 * with one FMA residual correction; range extremes keep the division. */
static double crossing_quotient(double delta, double direction, double inverse) {
    if (inverse == 0.0) return delta / direction;
    const double q = delta * inverse;
    if (!isfinite(q) || fabs(q) < 2.2250738585072014e-308) return delta / direction;
    return fma(fma(-q, direction, delta), inverse, q);
}

/* kernels/projection.py:643-655 */
static double cell_crossing(const ref_ray *r, int axis, int plane, const int *shape, double upper, double inverse) {
    if (r->direction[axis] == 0.0 || plane < 0 || plane >= shape[axis]) return upper;
    const double c = crossing_quotient((double)plane - r->origin[axis], r->direction[axis], inverse);
    return c < upper ? c : upper;
}

/* kernels/projection.py:659-695 (begin_cells) */
static cell_walk begin_cells(const ref_ray *r, const interval64 *iv, const int *shape) {
    cell_walk w;
    memset(&w, 0, sizeof w);
    for (int a = 0; a < 3; ++a) {
        const double entry = r->origin[a] + iv->lower * r->direction[a];
        const double v = r->direction[a];
        if (v != 0.0) {
            const double inverse = 1.0 / v;
            if (isfinite(inverse) && fabs(inverse) >= 2.2250738585072014e-308 && fabs(v) >= 2.2250738585072014e-308)
                w.inv[a] = inverse;
        }
        if (v > 0.0) {
            w.step[a] = 1;
            const int c = (int)ceil(entry);
            w.plane[a] = c > 0 ? c : 0;
        } else if (v < 0.0) {
            w.step[a] = -1;
            const int f = (int)floor(entry);
            w.plane[a] = f < shape[a] - 1 ? f : shape[a] - 1;
        }
        double crossing = cell_crossing(r, a, w.plane[a], shape, iv->upper, w.inv[a]);
        if (crossing <= iv->lower) {
            w.plane[a] += w.step[a];
            crossing = cell_crossing(r, a, w.plane[a], shape, iv->upper, w.inv[a]);
        }
        w.next[a] = crossing;
    }
    return w;
}

typedef struct {
    double end, radius, first_position, second_position;
    double first[3], second[3];
    int lower[3];
    int shared;
} cell_piece;

/* kernels/projection.py:710-742 (cell_segment) */
static cell_piece cell_segment(const ref_ray *r, const cell_walk *w, double start, const int *shape) {
    cell_piece s;
    const double m = w->next[1] < w->next[2] ? w->next[1] : w->next[2];
    s.end = w->next[0] < m ? w->next[0] : m;
    s.radius = (s.end - start) * 0.5;
    const double centre = start + s.radius;
    const double offset = s.radius / sqrt(3.0);
    s.first_position = centre - offset;
    s.second_position = centre + offset;
    double mid[3];
    for (int a = 0; a < 3; ++a) {
        s.first[a] = r->origin[a] + s.first_position * r->direction[a];
        s.second[a] = r->origin[a] + s.second_position * r->direction[a];
        mid[a] = r->origin[a] + centre * r->direction[a];
    }
    double unused[3];
    cell64(mid, shape, s.lower, unused);
    s.shared = 1;
    for (int a = 0; a < 3; ++a) {
        const double lo = (double)s.lower[a], up = lo + 1.0;
        const double c1 = clampd(s.first[a], 0.0, (double)(shape[a] - 1));
        const double c2 = clampd(s.second[a], 0.0, (double)(shape[a] - 1));
        if (c1 < lo || c2 < lo) s.shared = 0;
        if (s.lower[a] < shape[a] - 2 && (c1 >= up || c2 >= up)) s.shared = 0;
    }
    return s;
}

/* kernels/projection.py:745-755 (advance_cells): every tied plane advances */
static void advance_cells(cell_walk *w, const ref_ray *r, const int *shape, double end, double upper) {
    for (int a = 0; a < 3; ++a)
        if (w->next[a] <= end) {
            w->plane[a] += w->step[a];
            w->next[a] = cell_crossing(r, a, w->plane[a], shape, upper, w->inv[a]);
        }
}

/* sample_cell with a given cell (kernels/projection.py:278-284) or
 * sample_field (:288-295) when lower is NULL: value and slope, zero outside
 * the support. */
static double sample_point64(const float *field, const double *pt, const int *shape, const int *lower,
                             double *grad) {
    if (grad) grad[0] = grad[1] = grad[2] = 0.0;
    if (!inside64(pt, shape)) return 0.0;
    int own[3];
    double f[3], v[8];
    cell64(pt, shape, own, f);
    if (lower != NULL)
        for (int a = 0; a < 3; ++a) f[a] = clampd(pt[a], 0.0, (double)(shape[a] - 1)) - (double)lower[a];
    load_cell(field, lower != NULL ? lower : own, shape, v);
    return sample_cell64(v, pt, shape, f, grad);
}

/* Walk the cells of one clipped ray; returns the number of pieces and sets
 * BIT_RANGE on a stalled or unfinished walk (kernels/projection.py:714-745). */
typedef void (*piece_action)(const cell_piece *s, void *ctx);

static void walk_cells(const ref_ray *r, const interval64 *iv, const int *shape, int32_t *status, int view,
                       piece_action act, void *ctx) {
    cell_walk w = begin_cells(r, iv, shape);
    double start = iv->lower;
    long segments = 0;
    const long limit = (long)shape[0] + shape[1] + shape[2] + 1;
    while (start < iv->upper && segments < limit) {
        cell_piece s = cell_segment(r, &w, start, shape);
        if (s.end <= start) {
            status_or(status, view, BIT_RANGE);
            break;
        }
        if (!finite3(s.first) || !finite3(s.second)) status_or(status, view, BIT_RANGE);
        act(&s, ctx);
        start = s.end;
        advance_cells(&w, r, shape, s.end, iv->upper);
        ++segments;
    }
    if (start < iv->upper) status_or(status, view, BIT_RANGE);
}

typedef struct {
    const float *field;
    const int *shape;
    double integral;
} forward_ctx;

/* kernels/projection.py:696-711 (_scalar_segment_action) */
static void forward_piece(const cell_piece *s, void *vctx) {
    forward_ctx *c = (forward_ctx *)vctx;
    const int *lower = s->shared ? s->lower : NULL;
    c->integral += s->radius * sample_point64(c->field, s->first, c->shape, lower, NULL);
    c->integral += s->radius * sample_point64(c->field, s->second, c->shape, lower, NULL);
}

/* kernels/projection.py:802-846 with cell_gauss: integral * length */
void ref_forward_gauss_f64(const ref_plan *p, const float *field, const double *poses, double *out,
                           int32_t *status) {
    const int shape[3] = {p->nx, p->ny, p->nz};
    const int pixels = p->rows * p->cols;
    const long total_rays = (long)p->views * pixels;
#pragma omp parallel for schedule(dynamic, 64)
    for (long ray = 0; ray < total_rays; ++ray) {
        const int view = (int)(ray / pixels), pixel = (int)(ray % pixels);
        ref_ray r;
        ray_setup(p, poses + 12 * (size_t)view, view, pixel, &r);
        if (!valid_ray(&r)) {
            status_or(status, view, BIT_RANGE);
            out[ray] = 0.0;
            continue;
        }
        interval64 iv = finite_interval64(r.origin, r.direction, shape);
        double integral = 0.0;
        if (iv.upper > iv.lower) {
            forward_ctx c = {field, shape, 0.0};
            walk_cells(&r, &iv, shape, status, view, forward_piece, &c);
            integral = c.integral * r.length;
        }
        out[ray] = integral;
        if (!isfinite(integral)) status_or(status, view, BIT_RANGE);
    }
}

typedef struct {
    const float *field;
    const int *shape;
    int32_t *status;
    int view;
    double factor, seed, length;
    double *grad_field;
    int pose;
    double og[3], dg[3];
    double integral; /* sum of radius x node value */
} vjp_ctx;

/* kernels/projection.py:375-393 / 405-430 with FP32 increments, accumulated
 * in FP64 */
static void scatter_point(vjp_ctx *c, const double *pt, double vf) {
    if (!inside64(pt, c->shape)) return;
    int lower[3];
    double f[3];
    cell64(pt, c->shape, lower, f);
    for (int k = 0; k < 8; ++k) {
        double contribution = vf;
        for (int a = 0; a < 3; ++a) contribution *= ((k >> a) & 1) ? f[a] : 1.0 - f[a];
        const float rounded = (float)contribution;
        if (!isfinite(rounded)) status_or(c->status, c->view, BIT_RANGE);
        add_f64(&c->grad_field[corner_index(lower, c->shape, k)], (double)rounded);
    }
}

/* kernels/projection.py:936-1030 (cell-Gauss piece of the VJP) */
static void vjp_piece(const cell_piece *s, void *vctx) {
    vjp_ctx *c = (vjp_ctx *)vctx;
    if (c->pose) {
        double g1[3], g2[3];
        const int *lower = s->shared ? s->lower : NULL;
        const double v1 = sample_point64(c->field, s->first, c->shape, lower, g1);
        const double v2 = sample_point64(c->field, s->second, c->shape, lower, g2);
        c->integral += s->radius * v1;
        c->integral += s->radius * v2;
        for (int a = 0; a < 3; ++a) {
            c->og[a] += s->radius * g1[a];
            c->og[a] += s->radius * g2[a];
        }
        for (int a = 0; a < 3; ++a) {
            c->dg[a] += s->radius * s->first_position * g1[a];
            c->dg[a] += s->radius * s->second_position * g2[a];
        }
    }
    if (c->grad_field != NULL) {
        double vf = c->factor * s->radius;
        if (!isfinite(c->factor)) vf = product3(c->seed, c->length, s->radius);
        if (s->shared) {
            /* kernels/projection.py:405-430 (scatter_cell): both nodes' weights
             * summed in FP64, one FP32 rounding per corner */
            double f1[3], f2[3];
            for (int a = 0; a < 3; ++a) {
                f1[a] = clampd(s->first[a], 0.0, (double)(c->shape[a] - 1)) - (double)s->lower[a];
                f2[a] = clampd(s->second[a], 0.0, (double)(c->shape[a] - 1)) - (double)s->lower[a];
            }
            const double w1 = inside64(s->first, c->shape) ? vf : 0.0;
            const double w2 = inside64(s->second, c->shape) ? vf : 0.0;
            for (int k = 0; k < 8; ++k) {
                double first = w1, second = w2;
                for (int a = 0; a < 3; ++a) first *= ((k >> a) & 1) ? f1[a] : 1.0 - f1[a];
                for (int a = 0; a < 3; ++a) second *= ((k >> a) & 1) ? f2[a] : 1.0 - f2[a];
                const float rounded = (float)(first + second);
                if (!isfinite(rounded)) status_or(c->status, c->view, BIT_RANGE);
                add_f64(&c->grad_field[corner_index(s->lower, c->shape, k)], (double)rounded);
            }
        } else {
            scatter_point(c, s->first, vf);
            scatter_point(c, s->second, vf);
        }
    }
}

/* kernels/projection.py:873-1237 with cell_gauss (FP32 volume destination) */
void ref_vjp_gauss_f64(const ref_plan *p, const float *field, const double *poses, const double *seeds,
                       double *grad_field, double *grad_pose, double *per_ray_out, int components,
                       int32_t *status) {
    const int shape[3] = {p->nx, p->ny, p->nz};
    const int pixels = p->rows * p->cols;
    const long total_rays = (long)p->views * pixels;
    const int active_pose = grad_pose != NULL || per_ray_out != NULL;
    double *per_ray = per_ray_out;
    if (active_pose && per_ray == NULL)
        per_ray = (double *)calloc((size_t)total_rays * outputs_of(components), sizeof(double));
#pragma omp parallel for schedule(dynamic, 64)
    for (long ray = 0; ray < total_rays; ++ray) {
        const int view = (int)(ray / pixels), pixel = (int)(ray % pixels);
        const double *pose = poses + 12 * (size_t)view;
        double gradient[12] = {0};
        ref_ray r;
        ray_setup(p, pose, view, pixel, &r);
        if (!valid_ray(&r)) {
            status_or(status, view, BIT_RANGE);
        } else {
            interval64 iv = finite_interval64(r.origin, r.direction, shape);
            double span = iv.upper - iv.lower;
            const double seed = seeds[ray];
            if (seed == 0.0) span = 0.0;
            if (span > 0.0) {
                vjp_ctx c;
                memset(&c, 0, sizeof c);
                c.field = field;
                c.shape = shape;
                c.status = status;
                c.view = view;
                c.seed = seed;
                c.length = r.length;
                c.factor = seed * r.length;
                c.grad_field = grad_field;
                c.pose = active_pose;
                walk_cells(&r, &iv, shape, status, view, vjp_piece, &c);
                if (active_pose) {
                    /* kernels/projection.py:1031-1046: endpoint terms of the
                     * zero-extended support (support_value clamps the clipped
                     * endpoints into the support) */
                    double lo_pt[3], up_pt[3];
                    for (int a = 0; a < 3; ++a) {
                        lo_pt[a] = clampd(r.origin[a] + iv.lower * r.direction[a], -0.5, (double)shape[a] - 0.5);
                        up_pt[a] = clampd(r.origin[a] + iv.upper * r.direction[a], -0.5, (double)shape[a] - 0.5);
                    }
                    const double lower_value = sample_point64(field, lo_pt, shape, NULL, NULL);
                    const double upper_value = sample_point64(field, up_pt, shape, NULL, NULL);
                    for (int a = 0; a < 3; ++a) {
                        c.og[a] += upper_value * iv.up_o[a] - lower_value * iv.lo_o[a];
                        c.dg[a] += upper_value * iv.up_d[a] - lower_value * iv.lo_d[a];
                    }
                    if (isfinite(c.factor) && components == REF_RAY_COORDINATES) {
                        ray_epilogue(p, pose, &r, c.factor, c.og, c.dg, seed, c.integral, 1, gradient);
                    } else if (isfinite(c.factor)) {
                        pose_epilogue(p, pose, &r, c.factor, c.og, c.dg, components, gradient);
                    } else {
                        double sa[3], da[3];
                        for (int i = 0; i < 3; ++i) {
                            const double sg = p->g[i] * c.og[0] + p->g[3 + i] * c.og[1] + p->g[6 + i] * c.og[2];
                            const double rg = p->g[i] * c.dg[0] + p->g[3 + i] * c.dg[1] + p->g[6 + i] * c.dg[2];
                            sa[i] = product4(seed, r.length, 1.0, sg);
                            da[i] = product4(seed, r.length, 1.0, rg);
                        }
                        ref_plan unit = *p;
                        memset(unit.g, 0, sizeof unit.g);
                        unit.g[0] = unit.g[4] = unit.g[8] = 1.0;
                        if (components == REF_RAY_COORDINATES)
                            ray_epilogue(&unit, pose, &r, 1.0, sa, da, seed, c.integral, 1, gradient);
                        else
                            pose_epilogue(&unit, pose, &r, 1.0, sa, da, components, gradient);
                    }
                }
            }
        }
        if (active_pose)
            for (int k = 0; k < outputs_of(components); ++k)
                per_ray[(size_t)ray * outputs_of(components) + k] = gradient[k];
    }
    if (grad_pose != NULL) reduce_pose(p, per_ray, components, grad_pose);
    if (per_ray != per_ray_out) free(per_ray);
}
