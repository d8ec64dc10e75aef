/*
 * CPU restatement of the reference photon-transport walk (TEST INFRASTRUCTURE).
 *
 * This file is part of the oracle: only tests, the smoke check and the
 * CPU-baseline leg of bench.py may load it, and only as the checker.  The
 * product path (dpt) never calls into it.
 *
 * It restates, in binary64 and operation for operation, the analogue
 * Woodcock transport of the reference:
 *
 *   Philox4x32-10 random4          kernels/random.py:27-60
 *   point-cone source sampling     transport/kernels.py:2339-2366,
 *                                  transport/source.py:86-182
 *   spectrum index                 transport/spectrum_kernels.py:8-18
 *   detector_pixel                 transport/kernels.py:247-281
 *   grid_face / cell_index         transport/kernels.py:284-319
 *   energy_coordinate etc.         transport/kernels.py:322-384
 *   scatter_direction              transport/kernels.py:424-436
 *   Klein-Nishina rejection        transport/kernels.py:439-516
 *   enter_grid / faces / flight    transport/kernels.py:551-684
 *   walk (unrecorded)              transport/kernels.py:941-1563
 *   Rayleigh rejection             transport/physical_kernels.py:67-127
 *   trace_histories                transport/kernels.py:1944-1994
 *
 * Built with -ffp-contract=off and no fused operations: the reference module
 * sets fuse_fp=False for these kernels (transport/kernels.py:41).  The
 * readout, derivative, deposition and recording branches of the reference
 * walk are not part of the simulated path and are omitted; fluorescence is
 * supported only as "no lines" (absorption ends the history).
 *
 * Outputs per history follow trace_histories (pixel, score, energy, events,
 * status) plus the real-interaction count that separates primary (zero real
 * interactions) from scatter (transport/readout.py:19-24).  Tallies are the
 * per-pixel sum and sum of squares of the scores, split by that class.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum {
    ESCAPED = 0,
    ABSORBED = 1,
    EVENT_BUDGET = 2,
    CROSSING_BUDGET = 3,
    NUMERICAL_FAILURE = 4,
    ENERGY_SUPPORT = 5,
    ANGLE_BUDGET = 6,
};
#define MISS_PIXEL (-1)
#define INVALID_PIXEL (-2)
#define NAMESPACE_BIT 0x80000000u
static const double MAX_DISTANCE = 1.7976931348623157e308;
static const double TWO_PI = 6.2831853071795864769;
static const double ELECTRON_REST_KEV = 510.99895069;
static const double HC_KEV_ANGSTROM = 12.398419843320026;
static const double PI_D = 3.14159265358979323846;

/* Model, mirrored field for field by the product's dpt_transport_model. */
typedef struct {
    int32_t nx, ny, nz, materials, energy_bins, spectrum_bins;
    int32_t max_events, max_crossings, max_angle_trials;
    int32_t compton;        /* 0 isotropic elastic, 1 free-electron Klein-Nishina */
    int32_t compton_packed; /* packed two-proposal sampler */
    int32_t energy_score;   /* score photon energy instead of counts */
    int32_t physics;        /* Rayleigh channel / spectrum / majorant tables bound */
    int32_t woodcock;       /* majorant table present */
    int32_t det_shape[2];   /* width (columns), height (rows) */
    double origin[3], spacing[3];
    double det_corner[3], det_u[3], det_v[3], det_normal[3], det_spacing[2];
    double inv_rot[9], inv_trans[3];
    double source_energy, cutoff;
    const uint8_t *material_ids; /* x fastest */
    const double *energies;      /* K */
    const double *absorption;    /* M*K, energy fastest, mm^-1 at unit density */
    const double *scattering;    /* M*K */
    const double *rayleigh;      /* M*K (physics) */
    const double *majorant;      /* K (woodcock) */
    const double *majorant_max_density; /* M: 1 where the material occurs, else 0 */
    const int32_t *q_offsets;    /* M+1 */
    const double *q, *form_factor, *atomic_number;
    const double *spectrum_energies, *spectrum_cdf;
} tr_model;

typedef struct {
    double source[3], corner[3], u[3], v[3], extent[2], weight;
} tr_source;

/* ---------------------------------------------------------------- Philox */
/* kernels/random.py:27-60: counter (id lo, id hi, event, domain), key seed */
void tr_random4(uint64_t seed, uint64_t identity, uint32_t event, uint32_t domain, double out[4]) {
    uint32_t c0 = (uint32_t)identity, c1 = (uint32_t)(identity >> 32), c2 = event, c3 = domain;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    const double scale = 1.0 / 4294967296.0;
    out[0] = ((double)c0 + 0.5) * scale;
    out[1] = ((double)c1 + 0.5) * scale;
    out[2] = ((double)c2 + 0.5) * scale;
    out[3] = ((double)c3 + 0.5) * scale;
}

/* Raw words (for the known-answer test). */
void tr_philox_words(uint64_t seed, uint64_t identity, uint32_t event, uint32_t domain, uint32_t out[4]) {
    double u[4];
    tr_random4(seed, identity, event, domain, u);
    for (int i = 0; i < 4; ++i) out[i] = (uint32_t)(u[i] * 4294967296.0 - 0.5);
}

/* ------------------------------------------------------------ vec helpers */
static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void cross3(const double a[3], const double b[3], double r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* Warp normalize: v / length(v) when length > 1e-6 (else zero). */
static void normalize3(double v[3]) {
    const double l = sqrt(dot3(v, v));
    if (l > 1.0e-6) {
        v[0] = v[0] / l;
        v[1] = v[1] / l;
        v[2] = v[2] / l;
    } else {
        v[0] = v[1] = v[2] = 0.0;
    }
}

/* kernels/numerics.py:11-56: mantissa/exponent split product of three */
static double product3(double a, double b, double c) {
    if (!isfinite(a) || !isfinite(b) || !isfinite(c)) return a * b * c;
    if (a == 0.0 || b == 0.0 || c == 0.0) return 0.0;
    int ea, eb, ec;
    const double ma = frexp(a, &ea), mb = frexp(b, &eb), mc = frexp(c, &ec);
    return ldexp(ma * mb * mc, ea + eb + ec);
}

/* -------------------------------------------------------------- geometry */
/* transport/kernels.py:284-297: o + i*s with two roundings */
static double grid_face(double origin, double spacing, int index) {
    const double offset = (double)index * spacing;
    return origin + offset;
}

/* transport/kernels.py:299-319 */
static void cell_index(const tr_model *p, const double pos[3], const double dir[3], int cell[3]) {
    const int shape[3] = {p->nx, p->ny, p->nz};
    for (int a = 0; a < 3; ++a) {
        const double coordinate = (pos[a] - p->origin[a]) / p->spacing[a];
        int index = (int)floor(coordinate);
        double face = grid_face(p->origin[a], p->spacing[a], index);
        if (pos[a] < face)
            index -= 1;
        else if (pos[a] >= grid_face(p->origin[a], p->spacing[a], index + 1))
            index += 1;
        face = grid_face(p->origin[a], p->spacing[a], index);
        if (dir[a] < 0.0 && pos[a] == face) index -= 1;
        cell[a] = index < 0 ? 0 : (index > shape[a] - 1 ? shape[a] - 1 : index);
    }
}

/* transport/kernels.py:247-281 (general oriented detector branch) */
static int detector_pixel(const tr_model *p, const double pos[3], const double dir[3]) {
    int pixel = MISS_PIXEL;
    const double denominator = dot3(dir, p->det_normal);
    const double rel[3] = {p->det_corner[0] - pos[0], p->det_corner[1] - pos[1], p->det_corner[2] - pos[2]};
    const double numerator = dot3(rel, p->det_normal);
    if (denominator > 0.0) {
        const double distance = numerator / denominator;
        if (!isfinite(distance)) return INVALID_PIXEL;
        if (distance >= 0.0) {
            double offset[3];
            for (int i = 0; i < 3; ++i) offset[i] = (pos[i] + distance * dir[i]) - p->det_corner[i];
            const double x = dot3(offset, p->det_u);
            const double y = dot3(offset, p->det_v);
            if (!isfinite(x) || !isfinite(y)) return INVALID_PIXEL;
            if (x >= 0.0 && y >= 0.0 && x < (double)p->det_shape[0] * p->det_spacing[0] &&
                y < (double)p->det_shape[1] * p->det_spacing[1]) {
                int column = (int)floor(x / p->det_spacing[0]);
                int row = (int)floor(y / p->det_spacing[1]);
                if (column > p->det_shape[0] - 1) column = p->det_shape[0] - 1;
                if (row > p->det_shape[1] - 1) row = p->det_shape[1] - 1;
                pixel = row * p->det_shape[0] + column;
            }
        }
    }
    return pixel;
}

/* transport/kernels.py:322-339 */
static void energy_coordinate(const tr_model *p, double energy, int *lower_out, double *fraction_out) {
    if (p->energy_bins == 1) {
        *lower_out = 0;
        *fraction_out = 0.0;
        return;
    }
    int lower = 0, upper = p->energy_bins - 1;
    while (upper - lower > 1) {
        const int middle = (lower + upper) / 2;
        if (p->energies[middle] <= energy)
            lower = middle;
        else
            upper = middle;
    }
    *lower_out = lower;
    *fraction_out = (energy - p->energies[lower]) / (p->energies[upper] - p->energies[lower]);
}

/* physical_kernels.py:67-75 (energy_table_value); coefficients_at is the same law */
static double table_value(const tr_model *p, const double *values, int material, int lower, double fraction) {
    const int offset = material * p->energy_bins;
    if (p->energy_bins == 1) return values[offset];
    return (1.0 - fraction) * values[offset + lower] + fraction * values[offset + lower + 1];
}

/* transport/kernels.py:363-384 */
static int direct_flight_allowed(const tr_model *p, const double *density, int lower, double fraction,
                                 double majorant) {
    for (int m = 0; m < p->materials; ++m) {
        if (p->majorant_max_density[m] > 0.0) {
            const double a = table_value(p, p->absorption, m, lower, fraction);
            const double s = table_value(p, p->scattering, m, lower, fraction);
            const double rate = density[m] * (a + s);
            if (!isfinite(rate) || rate >= majorant) return 0;
        }
    }
    return 1;
}

/* transport/kernels.py:424-436 */
static void scatter_direction(double dir[3], double cosine, double azimuth) {
    double aux[3] = {0.0, 0.0, 1.0};
    if (fabs(dir[2]) > 0.9) {
        aux[0] = 1.0;
        aux[2] = 0.0;
    }
    double tangent[3], bitangent[3];
    cross3(aux, dir, tangent);
    normalize3(tangent);
    cross3(dir, tangent, bitangent);
    const double t = 1.0 - cosine * cosine;
    const double sine = sqrt(t > 0.0 ? t : 0.0);
    const double ca = cos(azimuth), sa = sin(azimuth);
    double out[3];
    for (int i = 0; i < 3; ++i) {
        const double mix = ca * tangent[i] + sa * bitangent[i];
        out[i] = cosine * dir[i] + sine * mix;
    }
    normalize3(out);
    dir[0] = out[0];
    dir[1] = out[1];
    dir[2] = out[2];
}

typedef struct {
    double cosine, azimuth, energy;
    int status;
} angular_t;

/* transport/kernels.py:439-472 (legacy per-proposal rejection) */
static angular_t compton_event(const tr_model *p, double energy, uint64_t seed, uint64_t history, uint32_t event) {
    angular_t r = {0.0, 0.0, energy, ANGLE_BUDGET};
    for (int trial = 0; trial < p->max_angle_trials; ++trial) {
        double u[4];
        tr_random4(seed, history, event, NAMESPACE_BIT + (uint32_t)trial, u);
        const double cosine = 2.0 * u[0] - 1.0;
        const double ratio = 1.0 / (1.0 + energy / ELECTRON_REST_KEV * (1.0 - cosine));
        const double acceptance = 0.5 * (ratio * ratio * ratio + ratio - ratio * ratio * (1.0 - cosine * cosine));
        if (u[1] < acceptance) {
            r.cosine = cosine;
            r.azimuth = TWO_PI * u[2];
            r.energy = energy * ratio;
            r.status = ESCAPED;
            break;
        }
    }
    return r;
}

/* transport/kernels.py:475-516 (two proposals per block, azimuth from the collision draw) */
static angular_t compton_packed(const tr_model *p, double energy, uint64_t seed, uint64_t history, uint32_t event,
                                double azimuth_uniform) {
    angular_t r = {0.0, 0.0, energy, ANGLE_BUDGET};
    const double alpha = energy / ELECTRON_REST_KEV;
    for (int block = 0; block < (p->max_angle_trials + 1) / 2; ++block) {
        double u[4];
        tr_random4(seed, history, event, NAMESPACE_BIT + (uint32_t)block, u);
        for (int member = 0; member < 2; ++member) {
            if (2 * block + member < p->max_angle_trials && r.status != ESCAPED) {
                const double cosine = 2.0 * u[2 * member] - 1.0;
                const double ratio = 1.0 / (1.0 + alpha * (1.0 - cosine));
                const double acceptance =
                    0.5 * (ratio * ratio * ratio + ratio - ratio * ratio * (1.0 - cosine * cosine));
                if (u[2 * member + 1] < acceptance) {
                    r.cosine = cosine;
                    r.azimuth = TWO_PI * azimuth_uniform;
                    r.energy = energy * ratio;
                    r.status = ESCAPED;
                }
            }
        }
        if (r.status == ESCAPED) break;
    }
    return r;
}

/* physical_kernels.py:78-96 (coherent form factor, linear in q) */
static double form_factor(const tr_model *p, double q, int material) {
    int lower = p->q_offsets[material];
    int upper = p->q_offsets[material + 1] - 1;
    while (upper - lower > 1) {
        const int middle = (lower + upper) / 2;
        if (p->q[middle] <= q)
            lower = middle;
        else
            upper = middle;
    }
    const double fraction = (q - p->q[lower]) / (p->q[upper] - p->q[lower]);
    return (1.0 - fraction) * p->form_factor[lower] + fraction * p->form_factor[upper];
}

/* physical_kernels.py:99-127 */
static angular_t rayleigh_event(const tr_model *p, double energy, int material, uint64_t seed, uint64_t history,
                                uint32_t event) {
    angular_t r = {0.0, 0.0, energy, ANGLE_BUDGET};
    for (int trial = 0; trial < p->max_angle_trials; ++trial) {
        double u[4];
        tr_random4(seed, history, event, 0xC0000000u + (uint32_t)trial, u);
        const double cosine = 2.0 * u[0] - 1.0;
        const double q = energy / HC_KEV_ANGSTROM * sqrt((1.0 - cosine) * 0.5);
        const double factor = form_factor(p, q, material) / p->atomic_number[material];
        const double acceptance = 0.5 * (1.0 + cosine * cosine) * factor * factor;
        if (u[1] < acceptance) {
            r.cosine = cosine;
            r.azimuth = 2.0 * PI_D * u[2];
            r.status = ESCAPED;
            break;
        }
    }
    return r;
}

typedef struct {
    int pixel, events, real_events, status;
    double energy, score;
} trace_t;

static int flat_index(const tr_model *p, const int cell[3]) {
    return (cell[2] * p->ny + cell[1]) * p->nx + cell[0];
}

/* transport/kernels.py:941-1563 (walk, unrecorded, no readout) */
static trace_t walk(const tr_model *p, const double *density, const double init_pos[3], const double init_dir[3],
                    double weight, double amplitude, uint64_t seed, uint64_t history) {
    trace_t res = {MISS_PIXEL, 0, 0, ESCAPED, p->source_energy, 0.0};
    if (p->physics && p->spectrum_bins != 0) {
        double u[4];
        tr_random4(seed, history, 0x80000002u, 0u, u);
        int low = 0, high = p->spectrum_bins - 1;
        while (low < high) {
            const int mid = (low + high) / 2;
            if (u[0] < p->spectrum_cdf[mid])
                high = mid;
            else
                low = mid + 1;
        }
        res.energy = p->spectrum_energies[low];
    }
    double pos[3], dir[3];
    for (int i = 0; i < 3; ++i) {
        pos[i] = p->inv_rot[3 * i] * init_pos[0] + p->inv_rot[3 * i + 1] * init_pos[1] +
                 p->inv_rot[3 * i + 2] * init_pos[2] + p->inv_trans[i];
        dir[i] = p->inv_rot[3 * i] * init_dir[0] + p->inv_rot[3 * i + 1] * init_dir[1] +
                 p->inv_rot[3 * i + 2] * init_dir[2];
    }
    const int shape[3] = {p->nx, p->ny, p->nz};
    /* enter_grid, transport/kernels.py:551-592 */
    int cell[3] = {0, 0, 0};
    int alive;
    {
        double entry = 0.0, exit_distance = MAX_DISTANCE;
        int intersects = 1;
        for (int a = 0; a < 3; ++a) {
            const double lower = p->origin[a];
            const double upper = grid_face(lower, p->spacing[a], shape[a]);
            if (dir[a] == 0.0) {
                if (pos[a] < lower || pos[a] >= upper) intersects = 0;
            } else {
                const double first = (lower - pos[a]) / dir[a];
                const double second = (upper - pos[a]) / dir[a];
                if (!isfinite(first) || !isfinite(second)) {
                    res.status = NUMERICAL_FAILURE;
                    intersects = 0;
                    break;
                }
                const double near = first < second ? first : second;
                const double far = first > second ? first : second;
                entry = entry > near ? entry : near;
                exit_distance = exit_distance < far ? exit_distance : far;
            }
        }
        if (exit_distance <= entry) intersects = 0;
        alive = intersects;
        if (alive) {
            for (int a = 0; a < 3; ++a) {
                pos[a] = pos[a] + entry * dir[a];
                const double lo = p->origin[a], hi = grid_face(p->origin[a], p->spacing[a], shape[a]);
                pos[a] = pos[a] < lo ? lo : (pos[a] > hi ? hi : pos[a]);
            }
            cell_index(p, pos, dir, cell);
        }
    }
    int crossings = 0;
    const int allow_direct = p->physics && p->woodcock;
    int direct = 0;
    double checked_energy = -1.0;
    while (alive) {
        if (p->physics && res.energy <= p->cutoff) {
            res.energy = 0.0;
            res.status = ABSORBED;
            break;
        }
        if (res.events >= p->max_events) {
            res.status = EVENT_BUDGET;
            break;
        }
        if (res.energy < p->energies[0] || res.energy > p->energies[p->energy_bins - 1]) {
            res.status = ENERGY_SUPPORT;
            break;
        }
        double draws[4];
        tr_random4(seed, history, (uint32_t)res.events, 0u, draws);
        double residual = -log(draws[0]);
        int lower;
        double fraction;
        energy_coordinate(p, res.energy, &lower, &fraction);
        int collision = 0, material = 0, flat = 0;
        double a_rate = 0.0, s_rate = 0.0, physical_extinction = 0.0, extinction = 0.0, majorant = 0.0;
        if (p->physics && p->woodcock) majorant = table_value(p, p->majorant, 0, lower, fraction);
        if (allow_direct && checked_energy != res.energy) {
            direct = direct_flight_allowed(p, density, lower, fraction, majorant);
            checked_energy = res.energy;
        }
        if (direct) {
            /* woodcock_flight, transport/kernels.py:644-684 */
            if (!isfinite(majorant) || majorant <= 0.0) {
                alive = 0;
                res.status = NUMERICAL_FAILURE;
            } else {
                double distances[3] = {MAX_DISTANCE, MAX_DISTANCE, MAX_DISTANCE}, faces[3] = {0.0, 0.0, 0.0};
                for (int a = 0; a < 3; ++a) {
                    if (dir[a] != 0.0) {
                        double face = p->origin[a];
                        if (dir[a] > 0.0) face = grid_face(p->origin[a], p->spacing[a], shape[a]);
                        faces[a] = face;
                        const double d = (face - pos[a]) / dir[a];
                        distances[a] = d > 0.0 ? d : 0.0;
                    }
                }
                double exit_distance = distances[1] < distances[2] ? distances[1] : distances[2];
                exit_distance = distances[0] < exit_distance ? distances[0] : exit_distance;
                if (!isfinite(exit_distance)) {
                    alive = 0;
                    res.status = NUMERICAL_FAILURE;
                } else {
                    const double distance = residual / majorant;
                    if (distance < exit_distance) {
                        for (int a = 0; a < 3; ++a) pos[a] = pos[a] + distance * dir[a];
                        cell_index(p, pos, dir, cell);
                        alive = 1;
                    } else {
                        for (int a = 0; a < 3; ++a) pos[a] = pos[a] + exit_distance * dir[a];
                        for (int a = 0; a < 3; ++a)
                            if (dir[a] != 0.0 && distances[a] == exit_distance) pos[a] = faces[a];
                        alive = 0;
                        res.status = ESCAPED;
                    }
                }
            }
        }
        while (alive && !collision) {
            flat = flat_index(p, cell);
            material = p->material_ids[flat];
            a_rate = table_value(p, p->absorption, material, lower, fraction);
            s_rate = table_value(p, p->scattering, material, lower, fraction);
            extinction = density[material] * (a_rate + s_rate);
            physical_extinction = extinction;
            if (p->physics && p->woodcock) {
                extinction = majorant;
                if (physical_extinction >= extinction) {
                    res.status = NUMERICAL_FAILURE;
                    alive = 0;
                    break;
                }
            }
            if (direct) {
                collision = 1;
                break;
            }
            /* distances_to_faces, transport/kernels.py:594-611 */
            double fd[3] = {MAX_DISTANCE, MAX_DISTANCE, MAX_DISTANCE};
            for (int a = 0; a < 3; ++a) {
                if (dir[a] != 0.0) {
                    const int face_index = cell[a] + (dir[a] > 0.0 ? 1 : 0);
                    const double face = grid_face(p->origin[a], p->spacing[a], face_index);
                    double d = (face - pos[a]) / dir[a];
                    if (isfinite(d)) d = d > 0.0 ? d : 0.0;
                    fd[a] = d;
                }
            }
            double distance = fd[1] < fd[2] ? fd[1] : fd[2];
            distance = fd[0] < distance ? fd[0] : distance;
            if (!isfinite(distance)) {
                res.status = NUMERICAL_FAILURE;
                alive = 0;
                break;
            }
            const double optical = extinction * distance;
            if (!isfinite(extinction) || !isfinite(optical)) {
                res.status = NUMERICAL_FAILURE;
                alive = 0;
                break;
            }
            if (extinction > 0.0 && residual < optical) {
                distance = residual / extinction;
                for (int a = 0; a < 3; ++a) pos[a] = pos[a] + distance * dir[a];
                collision = 1;
            } else {
                residual = residual - optical;
                for (int a = 0; a < 3; ++a) pos[a] = pos[a] + distance * dir[a];
                /* cross_faces, transport/kernels.py:614-641 */
                int inside = 1;
                for (int a = 0; a < 3; ++a) {
                    if (fd[a] == distance) {
                        if (dir[a] > 0.0) {
                            cell[a] += 1;
                            pos[a] = grid_face(p->origin[a], p->spacing[a], cell[a]);
                        } else {
                            pos[a] = grid_face(p->origin[a], p->spacing[a], cell[a]);
                            cell[a] -= 1;
                        }
                    }
                    if (cell[a] < 0 || cell[a] >= shape[a]) inside = 0;
                }
                alive = inside;
                crossings += 1;
                if (alive && crossings >= p->max_crossings) {
                    res.status = CROSSING_BUDGET;
                    alive = 0;
                }
            }
        }
        if (collision) {
            res.events += 1;
            if (p->physics && p->woodcock) {
                double mark[4];
                tr_random4(seed, history, (uint32_t)(res.events - 1), 1u, mark);
                if (mark[0] >= physical_extinction / extinction) collision = 0;
            }
        }
        if (collision) {
            res.real_events += 1;
            const double absorption_probability = a_rate / (a_rate + s_rate);
            if (draws[1] < absorption_probability) {
                res.status = ABSORBED;
                alive = 0;
                if (p->physics) res.energy = 0.0; /* no fluorescence lines: all energy local */
            } else {
                double cosine = 2.0 * draws[2] - 1.0;
                double azimuth = TWO_PI * draws[3];
                int coherent = 0;
                if (p->physics) {
                    const double rayleigh_rate = table_value(p, p->rayleigh, material, lower, fraction);
                    double channel[4];
                    tr_random4(seed, history, (uint32_t)(res.events - 1), 3u, channel);
                    if (rayleigh_rate > s_rate) {
                        res.status = NUMERICAL_FAILURE;
                        alive = 0;
                    } else if (channel[0] < rayleigh_rate / s_rate) {
                        coherent = 1;
                        const angular_t sampled =
                            rayleigh_event(p, res.energy, material, seed, history, (uint32_t)(res.events - 1));
                        cosine = sampled.cosine;
                        azimuth = sampled.azimuth;
                        if (sampled.status != ESCAPED) {
                            res.status = ANGLE_BUDGET;
                            alive = 0;
                        }
                    }
                }
                if (alive && p->compton && !coherent) {
                    const angular_t angular =
                        p->compton_packed
                            ? compton_packed(p, res.energy, seed, history, (uint32_t)(res.events - 1), draws[2])
                            : compton_event(p, res.energy, seed, history, (uint32_t)(res.events - 1));
                    cosine = angular.cosine;
                    azimuth = angular.azimuth;
                    res.energy = angular.energy;
                    if (angular.status != ESCAPED) {
                        res.status = ANGLE_BUDGET;
                        alive = 0;
                    }
                }
                if (alive) scatter_direction(dir, cosine, azimuth);
            }
        }
    }
    if (res.status == ESCAPED) {
        res.pixel = detector_pixel(p, pos, dir);
        if (res.pixel == INVALID_PIXEL) {
            res.pixel = MISS_PIXEL;
            res.status = NUMERICAL_FAILURE;
        }
        if (res.pixel >= 0) res.score = product3(weight, amplitude, p->energy_score ? res.energy : 1.0);
    }
    if (!isfinite(res.score)) res.status = NUMERICAL_FAILURE;
    return res;
}

/* transport/kernels.py:2339-2366 (sample_point_source) */
static void point_source(const tr_source *s, uint64_t seed, uint64_t history, double pos[3], double dir[3]) {
    double u[4];
    tr_random4(seed, history, NAMESPACE_BIT, 0u, u);
    const double t0 = s->extent[0] * u[0], t1 = s->extent[1] * u[1];
    double delta[3];
    for (int i = 0; i < 3; ++i) {
        const double endpoint = (s->corner[i] + t0 * s->u[i]) + t1 * s->v[i];
        pos[i] = s->source[i];
        delta[i] = endpoint - s->source[i];
    }
    double scale = fabs(delta[1]) > fabs(delta[2]) ? fabs(delta[1]) : fabs(delta[2]);
    scale = fabs(delta[0]) > scale ? fabs(delta[0]) : scale;
    for (int i = 0; i < 3; ++i) dir[i] = delta[i] / scale;
    normalize3(dir);
}

/*
 * Trace `count` histories starting at identity `first`.  Per-history arrays
 * may be NULL.  `tally` (4 x pixels: primary sum, primary sum of squares,
 * scatter sum, scatter sum of squares) and `histogram` (7 counts) accumulate.
 * Returns the OR of 1 << status over failed histories (status >= 2).
 */
int32_t tr_trace(const tr_model *p, const double *density, const tr_source *s, uint64_t seed, uint64_t first,
                 int64_t count, double amplitude, int32_t *pixel, double *score, double *energy, int32_t *events,
                 int32_t *real_events, int32_t *status, double *tally, int64_t *histogram) {
    const int64_t pixels = (int64_t)p->det_shape[0] * p->det_shape[1];
    int32_t bits = 0;
#pragma omp parallel reduction(| : bits)
    {
        double *local = NULL;
        int64_t hist[7] = {0, 0, 0, 0, 0, 0, 0};
        if (tally) local = (double *)calloc((size_t)(4 * pixels), sizeof(double));
#pragma omp for schedule(dynamic, 4096)
        for (int64_t i = 0; i < count; ++i) {
            const uint64_t h = first + (uint64_t)i;
            double pos[3], dir[3];
            point_source(s, seed, h, pos, dir);
            const trace_t t = walk(p, density, pos, dir, s->weight, amplitude, seed, h);
            if (pixel) pixel[i] = t.pixel;
            if (score) score[i] = t.score;
            if (energy) energy[i] = t.energy;
            if (events) events[i] = t.events;
            if (real_events) real_events[i] = t.real_events;
            if (status) status[i] = t.status;
            hist[t.status] += 1;
            if (t.status >= EVENT_BUDGET) bits |= 1 << t.status;
            if (local && t.pixel >= 0) {
                const int64_t base = (t.real_events > 0 ? 2 : 0) * pixels;
                local[base + t.pixel] += t.score;
                local[base + pixels + t.pixel] += t.score * t.score;
            }
        }
#pragma omp critical
        {
            if (local)
                for (int64_t k = 0; k < 4 * pixels; ++k) tally[k] += local[k];
            if (histogram)
                for (int k = 0; k < 7; ++k) histogram[k] += hist[k];
        }
        free(local);
    }
    return bits;
}

/* Single-function probes for the host tests. */
void tr_scatter_direction(double dir[3], double cosine, double azimuth) { scatter_direction(dir, cosine, azimuth); }

int32_t tr_detector_pixel(const tr_model *p, const double pos[3], const double dir[3]) {
    return detector_pixel(p, pos, dir);
}

void tr_cell_index(const tr_model *p, const double pos[3], const double dir[3], int32_t cell[3]) {
    int c[3];
    cell_index(p, pos, dir, c);
    for (int i = 0; i < 3; ++i) cell[i] = c[i];
}

/* Accepted Klein-Nishina / Rayleigh samples for the distribution tests:
 * returns the cosine of the accepted proposal, or NaN if the budget ran out. */
double tr_compton_cosine(const tr_model *p, double energy, uint64_t seed, uint64_t history, uint32_t event,
                         double *out_energy) {
    const angular_t a = compton_event(p, energy, seed, history, event);
    if (out_energy) *out_energy = a.energy;
    return a.status == ESCAPED ? a.cosine : NAN;
}

double tr_rayleigh_cosine(const tr_model *p, double energy, int material, uint64_t seed, uint64_t history,
                          uint32_t event) {
    const angular_t a = rayleigh_event(p, energy, material, seed, history, event);
    return a.status == ESCAPED ? a.cosine : NAN;
}

int tr_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
