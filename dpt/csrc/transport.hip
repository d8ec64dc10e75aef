// Monte Carlo photon transport: analogue Woodcock tracking on gfx950.
//
// One thread carries one photon history at a time; every step of the per-
// history event loop follows the reference walk (transport/kernels.py:941-1563,
// unrecorded, no readout branches) in binary64 with no contraction, and the
// random numbers are the reference's counter-addressed Philox4x32-10
// (kernels/random.py:27-60), so a history's path is a function of
// (seed, history identity) alone, whatever thread or order runs it.
//
// trace_kernel: persistent grid.  Each wave keeps 64 photons in flight and
//   refills a lane as soon as its photon terminates (history regeneration):
//   the wave grabs identities in chunks of 256 from a device counter, and new
//   photons are ranked over the idle lanes with a ballot.  Lanes therefore do
//   not idle behind the longest history of a fixed batch.  The material and
//   interaction tables (a few KB) are staged once per workgroup in LDS; the
//   label volume is stored as 8-bit (2 MB at 128^3: resident in every XCD's
//   L2).  Rayleigh angles, whose rejection loop runs ~170 trials per event
//   at 60 keV, are drawn by the whole wave for one photon at a time, 64
//   trials per round.  A detected photon adds its score and its square to
//   the primary or scatter tally with FP64 global atomics; terminal statuses
//   are counted in LDS and flushed once per workgroup.
// moments_kernel: mean and variance of the mean per pixel and class.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dpt_amd.h"

namespace dpt_mc {

constexpr int kBlock = 256;
constexpr uint64_t kChunk = 256;
constexpr int kMaxM = DPT_TRANSPORT_MAX_MATERIALS;
constexpr double kMaxDistance = 1.7976931348623157e308;
constexpr double kTwoPi = 6.2831853071795864769;
constexpr double kPi = 3.14159265358979323846;
constexpr double kElectronRestKeV = 510.99895069;
constexpr double kHcKeVAngstrom = 12.398419843320026;
constexpr uint32_t kNamespaceBit = 0x80000000u;

enum : int {
    ESCAPED = DPT_HISTORY_ESCAPED,
    ABSORBED = DPT_HISTORY_ABSORBED,
    EVENT_BUDGET = DPT_HISTORY_EVENT_BUDGET,
    CROSSING_BUDGET = DPT_HISTORY_CROSSING_BUDGET,
    NUMERICAL_FAILURE = DPT_HISTORY_NUMERICAL_FAILURE,
    ENERGY_SUPPORT = DPT_HISTORY_ENERGY_SUPPORT,
    ANGLE_BUDGET = DPT_HISTORY_ANGLE_BUDGET,
};
constexpr int kMissPixel = -1;
constexpr int kInvalidPixel = -2;

// Bound model (kernel argument).  Table offsets are in doubles within the
// packed table buffer that each workgroup stages in LDS.
struct Model {
    int nx, ny, nz, M, K, S;
    int max_events, max_crossings, max_trials;
    int compton, packed, energy_score, physics, woodcock;
    int det_w, det_h;
    int off_energy, off_abs, off_sca, off_ray, off_maj, off_q, off_ff, off_z, off_spec_e, off_spec_c;
    int table_doubles;
    int q_off[kMaxM + 1];
    double origin[3], spacing[3];
    double det_corner[3], det_u[3], det_v[3], det_normal[3], det_spacing[2];
    double inv_rot[9], inv_trans[3];
    double source_energy, cutoff;
    double max_density[kMaxM];
};

// One trace call (kernel argument).
struct Call {
    double density[kMaxM];
    double src[3], corner[3], su[3], sv[3], extent[2], weight, amplitude;
    uint64_t seed, first;
    int64_t count;
    const uint8_t *ids;
    const double *tables;
    int32_t *pixel;
    double *score;
    double *energy;
    int32_t *events;
    int32_t *real_events;
    int32_t *status;
    double *tally;
    unsigned long long *histogram;
    int *failure_bits;
    unsigned long long *counter;
};

// ------------------------------------------------------------------ Philox
// kernels/random.py:27-60: counter (id lo, id hi, event, domain), key = seed.
__device__ __forceinline__ void random4(uint64_t seed, uint64_t identity, uint32_t event, uint32_t domain,
                                        double u[4]) {
    uint32_t c0 = (uint32_t)identity, c1 = (uint32_t)(identity >> 32), c2 = event, c3 = domain;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 multiply each (v_mad_u64_u32), not separate hi/lo
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        c0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        c1 = (uint32_t)p1;
        c2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    constexpr double scale = 1.0 / 4294967296.0;
    u[0] = ((double)c0 + 0.5) * scale;
    u[1] = ((double)c1 + 0.5) * scale;
    u[2] = ((double)c2 + 0.5) * scale;
    u[3] = ((double)c3 + 0.5) * scale;
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// v / |v| when |v| > 1e-6, else zero (Warp's normalize).
__device__ __forceinline__ void normalize3(double v[3]) {
    const double l = sqrt(dot3(v, v));
    if (l > 1.0e-6) {
        v[0] = v[0] / l;
        v[1] = v[1] / l;
        v[2] = v[2] / l;
    } else {
        v[0] = v[1] = v[2] = 0.0;
    }
}

// kernels/numerics.py:11-56: product of three with one mantissa split each.
__device__ __forceinline__ double product3(double a, double b, double c) {
    if (!isfinite(a) || !isfinite(b) || !isfinite(c)) return a * b * c;
    if (a == 0.0 || b == 0.0 || c == 0.0) return 0.0;
    int ea, eb, ec;
    const double ma = frexp(a, &ea), mb = frexp(b, &eb), mc = frexp(c, &ec);
    return ldexp(ma * mb * mc, ea + eb + ec);
}

// transport/kernels.py:284-297: o + i*s, two roundings.
__device__ __forceinline__ double grid_face(double origin, double spacing, int index) {
    const double offset = (double)index * spacing;
    return origin + offset;
}

__device__ __forceinline__ int axis_count(const Model &m, int a) { return a == 0 ? m.nx : (a == 1 ? m.ny : m.nz); }

// transport/kernels.py:299-319: downstream cell of a point, exact faces.
__device__ __forceinline__ void cell_index(const Model &m, const double pos[3], const double dir[3], int cell[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double coordinate = (pos[a] - m.origin[a]) / m.spacing[a];
        int index = (int)floor(coordinate);
        double face = grid_face(m.origin[a], m.spacing[a], index);
        if (pos[a] < face)
            index -= 1;
        else if (pos[a] >= grid_face(m.origin[a], m.spacing[a], index + 1))
            index += 1;
        face = grid_face(m.origin[a], m.spacing[a], index);
        if (dir[a] < 0.0 && pos[a] == face) index -= 1;
        const int n = axis_count(m, a);
        cell[a] = index < 0 ? 0 : (index > n - 1 ? n - 1 : index);
    }
}

// transport/kernels.py:247-281 (oriented detector).
__device__ __forceinline__ int detector_pixel(const Model &m, const double pos[3], const double dir[3]) {
    int pixel = kMissPixel;
    const double denominator = dot3(dir, m.det_normal);
    const double rel[3] = {m.det_corner[0] - pos[0], m.det_corner[1] - pos[1], m.det_corner[2] - pos[2]};
    const double numerator = dot3(rel, m.det_normal);
    if (denominator > 0.0) {
        const double distance = numerator / denominator;
        if (!isfinite(distance)) return kInvalidPixel;
        if (distance >= 0.0) {
            double offset[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) offset[i] = (pos[i] + distance * dir[i]) - m.det_corner[i];
            const double x = dot3(offset, m.det_u);
            const double y = dot3(offset, m.det_v);
            if (!isfinite(x) || !isfinite(y)) return kInvalidPixel;
            if (x >= 0.0 && y >= 0.0 && x < (double)m.det_w * m.det_spacing[0] &&
                y < (double)m.det_h * m.det_spacing[1]) {
                const int column = min((int)floor(x / m.det_spacing[0]), m.det_w - 1);
                const int row = min((int)floor(y / m.det_spacing[1]), m.det_h - 1);
                pixel = row * m.det_w + column;
            }
        }
    }
    return pixel;
}

// transport/kernels.py:322-339 (binary search over the LDS energy nodes).
__device__ __forceinline__ void energy_coordinate(const Model &m, const double *T, double energy, int &lower,
                                                  double &fraction) {
    if (m.K == 1) {
        lower = 0;
        fraction = 0.0;
        return;
    }
    const double *e = T + m.off_energy;
    int lo = 0, hi = m.K - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (e[mid] <= energy)
            lo = mid;
        else
            hi = mid;
    }
    lower = lo;
    fraction = (energy - e[lo]) / (e[hi] - e[lo]);
}

// physical_kernels.py:67-75 / transport/kernels.py:341-361.
__device__ __forceinline__ double table_value(const Model &m, const double *values, int material, int lower,
                                              double fraction) {
    const int offset = material * m.K;
    if (m.K == 1) return values[offset];
    return (1.0 - fraction) * values[offset + lower] + fraction * values[offset + lower + 1];
}

// transport/kernels.py:424-436.
__device__ __forceinline__ void scatter_direction(double dir[3], double cosine, double azimuth) {
    double aux[3] = {0.0, 0.0, 1.0};
    if (fabs(dir[2]) > 0.9) {
        aux[0] = 1.0;
        aux[2] = 0.0;
    }
    double t[3] = {aux[1] * dir[2] - aux[2] * dir[1], aux[2] * dir[0] - aux[0] * dir[2],
                   aux[0] * dir[1] - aux[1] * dir[0]};
    normalize3(t);
    const double b[3] = {dir[1] * t[2] - dir[2] * t[1], dir[2] * t[0] - dir[0] * t[2], dir[0] * t[1] - dir[1] * t[0]};
    const double s2 = 1.0 - cosine * cosine;
    const double sine = sqrt(s2 > 0.0 ? s2 : 0.0);
    double sa, ca;
    sincos(azimuth, &sa, &ca);
    double out[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double mix = ca * t[i] + sa * b[i];
        out[i] = cosine * dir[i] + sine * mix;
    }
    normalize3(out);
    dir[0] = out[0];
    dir[1] = out[1];
    dir[2] = out[2];
}

struct Angular {
    double cosine, azimuth, energy;
    int status;
};

// transport/kernels.py:439-472 (one Philox block per proposal).
__device__ __forceinline__ Angular compton_event(const Model &m, double energy, uint64_t seed, uint64_t history,
                                                 uint32_t event) {
    Angular r{0.0, 0.0, energy, ANGLE_BUDGET};
    for (int trial = 0; trial < m.max_trials; ++trial) {
        double u[4];
        random4(seed, history, event, kNamespaceBit + (uint32_t)trial, u);
        const double cosine = 2.0 * u[0] - 1.0;
        const double ratio = 1.0 / (1.0 + energy / kElectronRestKeV * (1.0 - cosine));
        const double acceptance = 0.5 * (ratio * ratio * ratio + ratio - ratio * ratio * (1.0 - cosine * cosine));
        if (u[1] < acceptance) {
            r.cosine = cosine;
            r.azimuth = kTwoPi * u[2];
            r.energy = energy * ratio;
            r.status = ESCAPED;
            break;
        }
    }
    return r;
}

// transport/kernels.py:475-516 (two proposals per block).
__device__ __forceinline__ Angular compton_packed(const Model &m, double energy, uint64_t seed, uint64_t history,
                                                  uint32_t event, double azimuth_uniform) {
    Angular r{0.0, 0.0, energy, ANGLE_BUDGET};
    const double alpha = energy / kElectronRestKeV;
    for (int block = 0; block < (m.max_trials + 1) / 2; ++block) {
        double u[4];
        random4(seed, history, event, kNamespaceBit + (uint32_t)block, u);
#pragma unroll
        for (int member = 0; member < 2; ++member) {
            if (2 * block + member < m.max_trials && r.status != ESCAPED) {
                const double cosine = 2.0 * u[2 * member] - 1.0;
                const double ratio = 1.0 / (1.0 + alpha * (1.0 - cosine));
                const double acceptance =
                    0.5 * (ratio * ratio * ratio + ratio - ratio * ratio * (1.0 - cosine * cosine));
                if (u[2 * member + 1] < acceptance) {
                    r.cosine = cosine;
                    r.azimuth = kTwoPi * azimuth_uniform;
                    r.energy = energy * ratio;
                    r.status = ESCAPED;
                }
            }
        }
        if (r.status == ESCAPED) break;
    }
    return r;
}

// physical_kernels.py:78-96 (coherent branch).
__device__ __forceinline__ double form_factor(const Model &m, const double *T, double q, int material) {
    const double *qs = T + m.off_q;
    const double *fs = T + m.off_ff;
    int lo = m.q_off[material], hi = m.q_off[material + 1] - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (qs[mid] <= q)
            lo = mid;
        else
            hi = mid;
    }
    const double fraction = (q - qs[lo]) / (qs[hi] - qs[lo]);
    return (1.0 - fraction) * fs[lo] + fraction * fs[hi];
}

// physical_kernels.py:99-127: rejection with c = 2U - 1, q = E/hc sqrt((1-c)/2),
// accept U' < (1 + c^2)/2 (F(q)/Z)^2; trial t draws from domain 0xC0000000 + t.
// Acceptance is ~0.6 % at 60 keV (~170 trials per event), so one lane's loop
// would hold its whole wave for hundreds of iterations.  The wave evaluates
// 64 consecutive trials of one photon at a time instead (lane l takes trial
// base + l) and keeps the first accepted one: the same trial, and the same
// cosine and azimuth, as the sequential loop.  Called with every lane of the
// wave active; every argument is uniform.
__device__ __forceinline__ Angular rayleigh_event_wave(const Model &m, const double *T, double energy, int material,
                                                       uint64_t seed, uint64_t history, uint32_t event, int lane) {
    Angular r{0.0, 0.0, energy, ANGLE_BUDGET};
    const double z = T[m.off_z + material];
    for (int base = 0; base < m.max_trials; base += 64) {
        const int trial = base + lane;
        bool accept = false;
        double cosine = 0.0, azimuth = 0.0;
        if (trial < m.max_trials) {
            double u[4];
            random4(seed, history, event, 0xC0000000u + (uint32_t)trial, u);
            cosine = 2.0 * u[0] - 1.0;
            const double q = energy / kHcKeVAngstrom * sqrt((1.0 - cosine) * 0.5);
            const double factor = form_factor(m, T, q, material) / z;
            const double acceptance = 0.5 * (1.0 + cosine * cosine) * factor * factor;
            accept = u[1] < acceptance;
            azimuth = 2.0 * kPi * u[2];
        }
        const uint64_t hits = __ballot(accept);
        if (hits != 0ull) {
            const int first = __ffsll((unsigned long long)hits) - 1;
            r.cosine = __shfl(cosine, first, 64);
            r.azimuth = __shfl(azimuth, first, 64);
            r.status = ESCAPED;
            break;
        }
    }
    return r;
}

// A Rayleigh event left for the wave to sample (event_step returns before
// the angular draw; the photon's remaining work is finish_rayleigh).
struct PendingRayleigh {
    bool pending;
    int material;
    uint32_t event;
};

// State of the photon a lane carries.
struct Photon {
    double pos[3], dir[3];
    double energy, checked_energy, coord_energy;
    double fraction, majorant;  // energy coordinate cache (valid at coord_energy)
    int lower;
    uint64_t history;
    int cell[3];
    int events, real_events, status, crossings;
    bool alive, direct;
};

// Source sample, spectrum draw, world->object transform and grid entry
// (transport/kernels.py:2339-2366, :984-1003, :551-592).
__device__ __forceinline__ void start(const Model &m, const Call &c, const double *T, uint64_t history, Photon &ph) {
    ph.history = history;
    ph.events = 0;
    ph.real_events = 0;
    ph.status = ESCAPED;
    ph.crossings = 0;
    ph.direct = false;
    ph.checked_energy = -1.0;
    ph.coord_energy = -1.0;
    ph.energy = m.source_energy;
    double u[4];
    random4(c.seed, history, kNamespaceBit, 0u, u);
    const double t0 = c.extent[0] * u[0], t1 = c.extent[1] * u[1];
    double delta[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double endpoint = (c.corner[i] + t0 * c.su[i]) + t1 * c.sv[i];
        delta[i] = endpoint - c.src[i];
    }
    const double scale = fmax(fabs(delta[0]), fmax(fabs(delta[1]), fabs(delta[2])));
    double d[3] = {delta[0] / scale, delta[1] / scale, delta[2] / scale};
    normalize3(d);
    if (m.physics && m.S != 0) {
        random4(c.seed, history, 0x80000002u, 0u, u);
        const double *cdf = T + m.off_spec_c;
        int low = 0, high = m.S - 1;
        while (low < high) {
            const int mid = (low + high) >> 1;
            if (u[0] < cdf[mid])
                high = mid;
            else
                low = mid + 1;
        }
        ph.energy = T[m.off_spec_e + low];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ph.pos[i] = m.inv_rot[3 * i] * c.src[0] + m.inv_rot[3 * i + 1] * c.src[1] + m.inv_rot[3 * i + 2] * c.src[2] +
                    m.inv_trans[i];
        ph.dir[i] = m.inv_rot[3 * i] * d[0] + m.inv_rot[3 * i + 1] * d[1] + m.inv_rot[3 * i + 2] * d[2];
    }
    double entry = 0.0, exit_distance = kMaxDistance;
    bool intersects = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!intersects && ph.status == NUMERICAL_FAILURE) break;
        const double lower = m.origin[a];
        const double upper = grid_face(lower, m.spacing[a], axis_count(m, a));
        if (ph.dir[a] == 0.0) {
            if (ph.pos[a] < lower || ph.pos[a] >= upper) intersects = false;
        } else {
            const double first = (lower - ph.pos[a]) / ph.dir[a];
            const double second = (upper - ph.pos[a]) / ph.dir[a];
            if (!isfinite(first) || !isfinite(second)) {
                ph.status = NUMERICAL_FAILURE;
                intersects = false;
            } else {
                entry = fmax(entry, fmin(first, second));
                exit_distance = fmin(exit_distance, fmax(first, second));
            }
        }
    }
    if (exit_distance <= entry) intersects = false;
    ph.alive = intersects;
    ph.cell[0] = ph.cell[1] = ph.cell[2] = 0;
    if (ph.alive) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double p = ph.pos[a] + entry * ph.dir[a];
            const double lo = m.origin[a], hi = grid_face(m.origin[a], m.spacing[a], axis_count(m, a));
            ph.pos[a] = p < lo ? lo : (p > hi ? hi : p);
        }
        cell_index(m, ph.pos, ph.dir, ph.cell);
    }
}

// One iteration of the reference's `while alive` event loop.
__device__ __forceinline__ void event_step(const Model &m, const Call &c, const double *T, Photon &ph,
                                           PendingRayleigh &pend) {
    if (m.physics && ph.energy <= m.cutoff) {
        ph.energy = 0.0;
        ph.status = ABSORBED;
        ph.alive = false;
        return;
    }
    if (ph.events >= m.max_events) {
        ph.status = EVENT_BUDGET;
        ph.alive = false;
        return;
    }
    if (ph.energy < T[m.off_energy] || ph.energy > T[m.off_energy + m.K - 1]) {
        ph.status = ENERGY_SUPPORT;
        ph.alive = false;
        return;
    }
    double draws[4];
    random4(c.seed, ph.history, (uint32_t)ph.events, 0u, draws);
    double residual = -log(draws[0]);
    // Energy node, fraction and majorant are pure functions of the energy:
    // computed once per energy, reused by every null event at that energy.
    if (ph.coord_energy != ph.energy) {
        energy_coordinate(m, T, ph.energy, ph.lower, ph.fraction);
        ph.majorant = (m.physics && m.woodcock) ? table_value(m, T + m.off_maj, 0, ph.lower, ph.fraction) : 0.0;
        ph.coord_energy = ph.energy;
    }
    const int lower = ph.lower;
    const double fraction = ph.fraction;
    bool collision = false;
    int material = 0;
    double a_rate = 0.0, s_rate = 0.0, physical_extinction = 0.0, extinction = 0.0;
    const double majorant = ph.majorant;
    if (m.physics && m.woodcock && ph.checked_energy != ph.energy) {
        // transport/kernels.py:363-384
        bool allowed = true;
        for (int k = 0; k < m.M; ++k) {
            if (m.max_density[k] > 0.0) {
                const double a = table_value(m, T + m.off_abs, k, lower, fraction);
                const double s = table_value(m, T + m.off_sca, k, lower, fraction);
                const double rate = c.density[k] * (a + s);
                if (!isfinite(rate) || rate >= majorant) allowed = false;
            }
        }
        ph.direct = allowed;
        ph.checked_energy = ph.energy;
    }
    if (ph.direct) {
        // woodcock_flight, transport/kernels.py:644-684
        if (!isfinite(majorant) || majorant <= 0.0) {
            ph.alive = false;
            ph.status = NUMERICAL_FAILURE;
        } else {
            double distances[3] = {kMaxDistance, kMaxDistance, kMaxDistance}, faces[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (ph.dir[a] != 0.0) {
                    const double face =
                        ph.dir[a] > 0.0 ? grid_face(m.origin[a], m.spacing[a], axis_count(m, a)) : m.origin[a];
                    faces[a] = face;
                    distances[a] = fmax((face - ph.pos[a]) / ph.dir[a], 0.0);
                }
            }
            const double exit_distance = fmin(distances[0], fmin(distances[1], distances[2]));
            if (!isfinite(exit_distance)) {
                ph.alive = false;
                ph.status = NUMERICAL_FAILURE;
            } else {
                const double distance = residual / majorant;
                if (distance < exit_distance) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) ph.pos[a] = ph.pos[a] + distance * ph.dir[a];
                    cell_index(m, ph.pos, ph.dir, ph.cell);
                } else {
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        ph.pos[a] = ph.pos[a] + exit_distance * ph.dir[a];
                        if (ph.dir[a] != 0.0 && distances[a] == exit_distance) ph.pos[a] = faces[a];
                    }
                    ph.alive = false;
                    ph.status = ESCAPED;
                }
            }
        }
    }
    while (ph.alive && !collision) {
        const int flat = (ph.cell[2] * m.ny + ph.cell[1]) * m.nx + ph.cell[0];
        material = c.ids[flat];
        a_rate = table_value(m, T + m.off_abs, material, lower, fraction);
        s_rate = table_value(m, T + m.off_sca, material, lower, fraction);
        extinction = c.density[material] * (a_rate + s_rate);
        physical_extinction = extinction;
        if (m.physics && m.woodcock) {
            extinction = majorant;
            if (physical_extinction >= extinction) {
                ph.status = NUMERICAL_FAILURE;
                ph.alive = false;
                break;
            }
        }
        if (ph.direct) {
            collision = true;
            break;
        }
        // distances_to_faces / cross_faces, transport/kernels.py:594-641
        double fd[3] = {kMaxDistance, kMaxDistance, kMaxDistance};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (ph.dir[a] != 0.0) {
                const int face_index = ph.cell[a] + (ph.dir[a] > 0.0 ? 1 : 0);
                const double face = grid_face(m.origin[a], m.spacing[a], face_index);
                double d = (face - ph.pos[a]) / ph.dir[a];
                if (isfinite(d)) d = fmax(d, 0.0);
                fd[a] = d;
            }
        }
        double distance = fmin(fd[0], fmin(fd[1], fd[2]));
        if (!isfinite(distance)) {
            ph.status = NUMERICAL_FAILURE;
            ph.alive = false;
            break;
        }
        const double optical = extinction * distance;
        if (!isfinite(extinction) || !isfinite(optical)) {
            ph.status = NUMERICAL_FAILURE;
            ph.alive = false;
            break;
        }
        if (extinction > 0.0 && residual < optical) {
            distance = residual / extinction;
#pragma unroll
            for (int a = 0; a < 3; ++a) ph.pos[a] = ph.pos[a] + distance * ph.dir[a];
            collision = true;
        } else {
            residual = residual - optical;
            bool inside = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                ph.pos[a] = ph.pos[a] + distance * ph.dir[a];
                if (fd[a] == distance) {
                    if (ph.dir[a] > 0.0) {
                        ph.cell[a] += 1;
                        ph.pos[a] = grid_face(m.origin[a], m.spacing[a], ph.cell[a]);
                    } else {
                        ph.pos[a] = grid_face(m.origin[a], m.spacing[a], ph.cell[a]);
                        ph.cell[a] -= 1;
                    }
                }
                if (ph.cell[a] < 0 || ph.cell[a] >= axis_count(m, a)) inside = false;
            }
            ph.alive = inside;
            ph.crossings += 1;
            if (ph.alive && ph.crossings >= m.max_crossings) {
                ph.status = CROSSING_BUDGET;
                ph.alive = false;
            }
        }
    }
    if (collision) {
        ph.events += 1;
        if (m.physics && m.woodcock) {
            double mark[4];
            random4(c.seed, ph.history, (uint32_t)(ph.events - 1), 1u, mark);
            if (mark[0] >= physical_extinction / extinction) collision = false;
        }
    }
    if (!collision) return;
    ph.real_events += 1;
    const uint32_t event = (uint32_t)(ph.events - 1);
    if (draws[1] < a_rate / (a_rate + s_rate)) {
        ph.status = ABSORBED;
        ph.alive = false;
        if (m.physics) ph.energy = 0.0;  // no fluorescence lines: all energy local
        return;
    }
    double cosine = 2.0 * draws[2] - 1.0;
    double azimuth = kTwoPi * draws[3];
    bool coherent = false;
    if (m.physics) {
        const double rayleigh_rate = table_value(m, T + m.off_ray, material, lower, fraction);
        double channel[4];
        random4(c.seed, ph.history, event, 3u, channel);
        if (rayleigh_rate > s_rate) {
            ph.status = NUMERICAL_FAILURE;
            ph.alive = false;
        } else if (channel[0] < rayleigh_rate / s_rate) {
            // sampled by the whole wave (rayleigh_event_wave), then finish_rayleigh
            pend.pending = true;
            pend.material = material;
            pend.event = event;
            return;
        }
    }
    if (ph.alive && m.compton && !coherent) {
        const Angular a = m.packed ? compton_packed(m, ph.energy, c.seed, ph.history, event, draws[2])
                                   : compton_event(m, ph.energy, c.seed, ph.history, event);
        cosine = a.cosine;
        azimuth = a.azimuth;
        ph.energy = a.energy;
        if (a.status != ESCAPED) {
            ph.status = ANGLE_BUDGET;
            ph.alive = false;
        }
    }
    if (ph.alive) scatter_direction(ph.dir, cosine, azimuth);
}

// The rest of a Rayleigh event once its angles are drawn.
__device__ __forceinline__ void finish_rayleigh(Photon &ph, const Angular &r) {
    if (r.status != ESCAPED) {
        ph.status = ANGLE_BUDGET;
        ph.alive = false;
        return;
    }
    scatter_direction(ph.dir, r.cosine, r.azimuth);
}

__global__ void zero_counter_kernel(unsigned long long *counter) { *counter = 0ull; }

// WAVES: minimum waves per SIMD the register allocation must allow (2 fits
// the whole walk in registers; 4 trades a few spills for occupancy).
template <int WAVES>
__global__ __launch_bounds__(kBlock, WAVES) void trace_kernel(Model m, Call c) {
    extern __shared__ double lds[];
    __shared__ unsigned int hist[8];
    for (int i = threadIdx.x; i < m.table_doubles; i += kBlock) lds[i] = c.tables[i];
    if (threadIdx.x < 8) hist[threadIdx.x] = 0u;
    __syncthreads();
    const double *T = lds;
    const int lane = threadIdx.x & 63;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint64_t count = (uint64_t)c.count;
    uint64_t q_next = 0, q_end = 0;
    bool exhausted = false, has = false;
    Photon ph;
    ph.alive = false;
    for (;;) {
        // Refill idle lanes from this wave's queue of history identities.
        uint64_t need = __ballot(!has);
        while (need != 0ull && !exhausted) {
            if (q_next >= q_end) {
                unsigned long long base = 0ull;
                if (lane == 0) base = atomicAdd(c.counter, (unsigned long long)kChunk);
                base = __shfl(base, 0);
                if (base >= count) {
                    exhausted = true;
                    break;
                }
                q_next = base;
                q_end = min(base + kChunk, count);
            }
            const uint64_t avail = q_end - q_next;
            const uint64_t rank = (uint64_t)__popcll(need & below);
            if (!has && rank < avail) {
                start(m, c, T, c.first + q_next + rank, ph);
                has = true;
            }
            q_next += min((uint64_t)__popcll(need), avail);
            need = __ballot(!has);
        }
        if (__ballot(has) == 0ull) break;
        PendingRayleigh pend{false, 0, 0u};
        if (has && ph.alive) event_step(m, c, T, ph, pend);
        // Rayleigh draws, one photon at a time by the whole wave (all lanes
        // active here; the loop bound and the leader's values are uniform)
        for (uint64_t waiting = __ballot(pend.pending); waiting != 0ull; waiting &= waiting - 1ull) {
            const int leader = __ffsll((unsigned long long)waiting) - 1;
            const Angular r = rayleigh_event_wave(m, T, __shfl(ph.energy, leader, 64), __shfl(pend.material, leader, 64),
                                                  c.seed, __shfl(ph.history, leader, 64),
                                                  (uint32_t)__shfl((int)pend.event, leader, 64), lane);
            if (lane == leader) finish_rayleigh(ph, r);
        }
        if (has) {
            if (!ph.alive) {
                // Termination: detector hit and score (transport/kernels.py:1520-1536).
                int pixel = kMissPixel;
                double score = 0.0;
                if (ph.status == ESCAPED) {
                    pixel = detector_pixel(m, ph.pos, ph.dir);
                    if (pixel == kInvalidPixel) {
                        pixel = kMissPixel;
                        ph.status = NUMERICAL_FAILURE;
                    }
                    if (pixel >= 0) score = product3(c.weight, c.amplitude, m.energy_score ? ph.energy : 1.0);
                }
                if (!isfinite(score)) ph.status = NUMERICAL_FAILURE;
                const uint64_t i = ph.history - c.first;
                if (c.pixel) c.pixel[i] = pixel;
                if (c.score) c.score[i] = score;
                if (c.energy) c.energy[i] = ph.energy;
                if (c.events) c.events[i] = ph.events;
                if (c.real_events) c.real_events[i] = ph.real_events;
                if (c.status) c.status[i] = ph.status;
                if (c.tally && pixel >= 0) {
                    const int64_t pixels = (int64_t)m.det_w * m.det_h;
                    double *base = c.tally + (ph.real_events > 0 ? 2 * pixels : 0);
                    unsafeAtomicAdd(base + pixel, score);
                    unsafeAtomicAdd(base + pixels + pixel, score * score);
                }
                atomicAdd(&hist[ph.status], 1u);
                if (ph.status >= EVENT_BUDGET) atomicOr(c.failure_bits, 1 << ph.status);
                has = false;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 7 && hist[threadIdx.x] != 0u)
        atomicAdd(c.histogram + threadIdx.x, (unsigned long long)hist[threadIdx.x]);
}

// transport/simulation.py:61-66, per class: mean = S/N, var = (Q - S*mean)/(N(N-1)).
__global__ void moments_kernel(const double *tally, int64_t pixels, double n, double *mean, double *variance) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    const double sp = tally[p], qp = tally[pixels + p], ss = tally[2 * pixels + p], qs = tally[3 * pixels + p];
    const double s[3] = {sp + ss, sp, ss}, q[3] = {qp + qs, qp, qs};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double mu = s[k] / n;
        const double m2 = q[k] - s[k] * mu;
        mean[k * pixels + p] = mu;
        variance[k * pixels + p] = (m2 > 0.0 ? m2 : 0.0) / (n * (n - 1.0));
    }
}

}  // namespace dpt_mc

using namespace dpt_mc;

struct dpt_transport {
    int device;
    Model model;
    std::vector<double> host_tables;
    std::vector<double> absorption, scattering, majorant;  // host copies for the density check
    uint8_t *d_ids;
    double *d_tables;
    unsigned long long *d_counter;
    unsigned long long *d_histogram;
    int *d_bits;
    unsigned long long *h_histogram;  // pinned: 7 counts + bits
    int grid2, grid4;
    int waves;  // 2 or 4: trace_kernel variant
};

namespace {

thread_local std::string g_tr_error;

int tr_fail(int code, const std::string &msg) {
    g_tr_error = msg;
    return code;
}

bool finite_nonnegative(const double *v, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i]) || v[i] < 0.0) return false;
    return true;
}

struct DeviceGuard {
    int previous = -1;
    explicit DeviceGuard(int device) {
        (void)hipGetDevice(&previous);
        (void)hipSetDevice(device);
    }
    ~DeviceGuard() {
        if (previous >= 0) (void)hipSetDevice(previous);
    }
};

}  // namespace

extern "C" {

const char *dpt_transport_last_error(void) { return g_tr_error.c_str(); }

int dpt_transport_destroy(dpt_transport *t) {
    if (t == nullptr) return DPT_OK;
    DeviceGuard guard(t->device);
    (void)hipFree(t->d_ids);
    (void)hipFree(t->d_tables);
    (void)hipFree(t->d_counter);
    (void)hipFree(t->d_histogram);
    (void)hipFree(t->d_bits);
    (void)hipHostFree(t->h_histogram);
    delete t;
    return DPT_OK;
}

int dpt_transport_create(const dpt_transport_model *md, int32_t device, dpt_transport **out) {
    if (out == nullptr || md == nullptr) return tr_fail(DPT_E_INVALID, "model and output handle are required");
    *out = nullptr;
    const dpt_transport_model &x = *md;
    if (x.nx < 1 || x.ny < 1 || x.nz < 1 || (int64_t)x.nx * x.ny * x.nz >= INT32_MAX)
        return tr_fail(DPT_E_INVALID, "grid extents must be positive and fit signed 32-bit indexing");
    if (x.materials < 1 || x.materials > kMaxM)
        return tr_fail(DPT_E_INVALID, "transport supports 1.." + std::to_string(kMaxM) + " materials");
    if (x.energy_bins < 1) return tr_fail(DPT_E_INVALID, "at least one coefficient energy node is required");
    if (x.max_events < 1 || x.max_crossings < 1 || x.max_angle_trials < 1)
        return tr_fail(DPT_E_INVALID, "event, crossing and angle budgets must be positive");
    if (x.det_shape[0] < 1 || x.det_shape[1] < 1 || (int64_t)x.det_shape[0] * x.det_shape[1] >= INT32_MAX)
        return tr_fail(DPT_E_INVALID, "detector extents must be positive");
    if (x.woodcock && !x.physics) return tr_fail(DPT_E_INVALID, "a Woodcock majorant requires physical tables");
    if (x.physics && x.max_angle_trials > 65536)
        return tr_fail(DPT_E_INVALID, "physical angular rejection reserves at most 65536 trial domains");
    for (int a = 0; a < 3; ++a)
        if (!(x.spacing[a] > 0.0) || !std::isfinite(x.spacing[a]) || !std::isfinite(x.origin[a]))
            return tr_fail(DPT_E_INVALID, "cell spacing must be positive and finite");
    if (x.material_ids == nullptr || x.energies == nullptr || x.absorption == nullptr || x.scattering == nullptr)
        return tr_fail(DPT_E_INVALID, "labels, energies, absorption and scattering are required");
    const int M = x.materials, K = x.energy_bins;
    for (int k = 1; k < K; ++k)
        if (!(x.energies[k] > x.energies[k - 1])) return tr_fail(DPT_E_INVALID, "coefficient energies must increase strictly");
    if (!finite_nonnegative(x.energies, K) || !finite_nonnegative(x.absorption, (int64_t)M * K) ||
        !finite_nonnegative(x.scattering, (int64_t)M * K))
        return tr_fail(DPT_E_INVALID, "coefficient tables must be finite and nonnegative");
    if (!(x.source_energy >= x.energies[0] && x.source_energy <= x.energies[K - 1]))
        return tr_fail(DPT_E_INVALID, "coefficient energies must contain the source energy");
    if (x.compton && K < 2) return tr_fail(DPT_E_INVALID, "Compton scattering needs energy-dependent tables");
    const int64_t voxels = (int64_t)x.nx * x.ny * x.nz;
    std::vector<double> present(M, 0.0);
    for (int64_t i = 0; i < voxels; ++i) {
        if (x.material_ids[i] >= M) return tr_fail(DPT_E_INVALID, "material id outside the material table");
        present[x.material_ids[i]] = 1.0;
    }
    int q_total = 0;
    if (x.physics) {
        if (x.rayleigh == nullptr || x.q_offsets == nullptr || x.q == nullptr || x.form_factor == nullptr ||
            x.atomic_number == nullptr)
            return tr_fail(DPT_E_INVALID, "physical transport requires Rayleigh and form-factor tables");
        if (!finite_nonnegative(x.rayleigh, (int64_t)M * K))
            return tr_fail(DPT_E_INVALID, "Rayleigh coefficients must be finite and nonnegative");
        for (int64_t i = 0; i < (int64_t)M * K; ++i)
            if (x.rayleigh[i] > x.scattering[i])
                return tr_fail(DPT_E_INVALID, "the Rayleigh partial exceeds aggregate scattering");
        if (x.q_offsets[0] != 0) return tr_fail(DPT_E_INVALID, "form-factor offsets must start at zero");
        double e_max = x.source_energy;
        if (x.spectrum_bins > 0) e_max = std::max(e_max, x.spectrum_energies[x.spectrum_bins - 1]);
        for (int k = 0; k < M; ++k) {
            const int lo = x.q_offsets[k], hi = x.q_offsets[k + 1];
            if (hi - lo < 2) return tr_fail(DPT_E_INVALID, "each form-factor table needs at least two nodes");
            if (x.q[lo] != 0.0) return tr_fail(DPT_E_INVALID, "form-factor q nodes must start at zero");
            for (int i = lo + 1; i < hi; ++i)
                if (!(x.q[i] > x.q[i - 1])) return tr_fail(DPT_E_INVALID, "form-factor q nodes must increase strictly");
            if (x.q[hi - 1] < e_max / kHcKeVAngstrom)
                return tr_fail(DPT_E_INVALID, "atomic scattering q support does not cover all source angles");
            if (!(x.atomic_number[k] >= 1.0) || x.form_factor[lo] != x.atomic_number[k])
                return tr_fail(DPT_E_INVALID, "atomic scattering requires F(0)=Z");
            for (int i = lo; i < hi; ++i)
                if (!(x.form_factor[i] >= 0.0) || x.form_factor[i] > x.atomic_number[k])
                    return tr_fail(DPT_E_INVALID, "form factor must lie in [0, Z]");
        }
        q_total = x.q_offsets[M];
        if (x.woodcock) {
            if (x.majorant == nullptr) return tr_fail(DPT_E_INVALID, "Woodcock transport requires a majorant table");
            for (int k = 0; k < K; ++k)
                if (!(x.majorant[k] > 0.0) || !std::isfinite(x.majorant[k]))
                    return tr_fail(DPT_E_INVALID, "Woodcock majorant must be positive and finite");
        }
        if (x.spectrum_bins > 0) {
            if (x.spectrum_energies == nullptr || x.spectrum_cdf == nullptr)
                return tr_fail(DPT_E_INVALID, "spectrum energies and CDF are required");
            double previous = 0.0;
            for (int s = 0; s < x.spectrum_bins; ++s) {
                if (!(x.spectrum_cdf[s] > previous)) return tr_fail(DPT_E_INVALID, "spectrum CDF must increase strictly");
                previous = x.spectrum_cdf[s];
                if (x.spectrum_energies[s] < x.energies[0] || x.spectrum_energies[s] > x.energies[K - 1])
                    return tr_fail(DPT_E_INVALID, "source spectrum leaves coefficient energy support");
            }
            if (previous != 1.0) return tr_fail(DPT_E_INVALID, "spectrum CDF must end at exactly one");
        }
    }
    auto *t = new dpt_transport();
    std::memset(&t->model, 0, sizeof t->model);
    Model &m = t->model;
    m.nx = x.nx;
    m.ny = x.ny;
    m.nz = x.nz;
    m.M = M;
    m.K = K;
    m.S = x.physics ? x.spectrum_bins : 0;
    m.max_events = x.max_events;
    m.max_crossings = x.max_crossings;
    m.max_trials = x.max_angle_trials;
    m.compton = x.compton;
    m.packed = x.compton_packed;
    m.energy_score = x.energy_score;
    m.physics = x.physics;
    m.woodcock = x.woodcock;
    m.det_w = x.det_shape[0];
    m.det_h = x.det_shape[1];
    std::memcpy(m.origin, x.origin, sizeof m.origin);
    std::memcpy(m.spacing, x.spacing, sizeof m.spacing);
    std::memcpy(m.det_corner, x.det_corner, sizeof m.det_corner);
    std::memcpy(m.det_u, x.det_u, sizeof m.det_u);
    std::memcpy(m.det_v, x.det_v, sizeof m.det_v);
    std::memcpy(m.det_normal, x.det_normal, sizeof m.det_normal);
    std::memcpy(m.det_spacing, x.det_spacing, sizeof m.det_spacing);
    std::memcpy(m.inv_rot, x.inv_rot, sizeof m.inv_rot);
    std::memcpy(m.inv_trans, x.inv_trans, sizeof m.inv_trans);
    m.source_energy = x.source_energy;
    m.cutoff = x.cutoff;
    for (int k = 0; k < M; ++k) m.max_density[k] = present[k];
    // packed tables
    std::vector<double> &tab = t->host_tables;
    auto append = [&](const double *v, int64_t n) {
        const int off = (int)tab.size();
        if (v != nullptr) tab.insert(tab.end(), v, v + n);
        return off;
    };
    m.off_energy = append(x.energies, K);
    m.off_abs = append(x.absorption, (int64_t)M * K);
    m.off_sca = append(x.scattering, (int64_t)M * K);
    m.off_ray = append(x.physics ? x.rayleigh : nullptr, x.physics ? (int64_t)M * K : 0);
    m.off_maj = append(x.woodcock ? x.majorant : nullptr, x.woodcock ? K : 0);
    m.off_q = append(x.physics ? x.q : nullptr, q_total);
    m.off_ff = append(x.physics ? x.form_factor : nullptr, q_total);
    m.off_z = append(x.physics ? x.atomic_number : nullptr, x.physics ? M : 0);
    m.off_spec_e = append(m.S ? x.spectrum_energies : nullptr, m.S);
    m.off_spec_c = append(m.S ? x.spectrum_cdf : nullptr, m.S);
    if (x.physics)
        for (int k = 0; k <= M; ++k) m.q_off[k] = x.q_offsets[k];
    m.table_doubles = (int)tab.size();
    if (tab.size() * sizeof(double) > DPT_TRANSPORT_MAX_TABLE_BYTES) {
        delete t;
        return tr_fail(DPT_E_INVALID, "transport tables exceed the " + std::to_string(DPT_TRANSPORT_MAX_TABLE_BYTES) +
                                          "-byte LDS budget");
    }
    t->absorption.assign(x.absorption, x.absorption + (int64_t)M * K);
    t->scattering.assign(x.scattering, x.scattering + (int64_t)M * K);
    if (x.woodcock) t->majorant.assign(x.majorant, x.majorant + K);
    t->device = device;
    DeviceGuard guard(device);
    if (hipMalloc(&t->d_ids, (size_t)voxels) != hipSuccess ||
        hipMalloc(&t->d_tables, std::max<size_t>(tab.size(), 1) * sizeof(double)) != hipSuccess ||
        hipMalloc(&t->d_counter, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&t->d_histogram, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&t->d_bits, sizeof(int)) != hipSuccess ||
        hipHostMalloc(&t->h_histogram, 8 * sizeof(unsigned long long)) != hipSuccess) {
        dpt_transport_destroy(t);
        return tr_fail(DPT_E_MEMORY, "device allocation failed");
    }
    if (hipMemcpy(t->d_ids, x.material_ids, (size_t)voxels, hipMemcpyHostToDevice) != hipSuccess ||
        (!tab.empty() &&
         hipMemcpy(t->d_tables, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemset(t->d_histogram, 0, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(t->d_bits, 0, sizeof(int)) != hipSuccess) {
        dpt_transport_destroy(t);
        return tr_fail(DPT_E_DEVICE, "device upload failed");
    }
    int per_cu2 = 0, per_cu4 = 0, cus = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) cus = prop.multiProcessorCount;
    const size_t lds = (size_t)m.table_doubles * sizeof(double);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, trace_kernel<2>, kBlock, lds) != hipSuccess) per_cu2 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu4, trace_kernel<4>, kBlock, lds) != hipSuccess) per_cu4 = 0;
    t->grid2 = std::max(1, per_cu2) * std::max(1, cus);
    t->grid4 = std::max(1, per_cu4) * std::max(1, cus);
    t->waves = 4;
    if (const char *env = std::getenv("DPT_AMD_MC_WAVES")) t->waves = std::atoi(env) == 2 ? 2 : 4;
    *out = t;
    return DPT_OK;
}

int dpt_transport_trace(dpt_transport *t, const double *density, const dpt_point_source *source, uint64_t seed,
                        uint64_t first_history, int64_t count, double amplitude, const dpt_history_records *records,
                        double *tally, void *stream) {
    if (t == nullptr || density == nullptr || source == nullptr)
        return tr_fail(DPT_E_INVALID, "transport handle, density and source are required");
    if (count < 1 || count >= ((int64_t)1 << 31)) return tr_fail(DPT_E_INVALID, "history count must be in 1..2^31-1");
    if (first_history + (uint64_t)count < first_history)
        return tr_fail(DPT_E_INVALID, "the history range would wrap its 64-bit identity");
    if (!std::isfinite(amplitude) || amplitude < 0.0 || !std::isfinite(source->weight) || source->weight < 0.0)
        return tr_fail(DPT_E_INVALID, "source weight and amplitude must be finite and nonnegative");
    const Model &m = t->model;
    for (int k = 0; k < m.M; ++k)
        if (!std::isfinite(density[k]) || !(density[k] > 0.0))
            return tr_fail(DPT_E_INVALID, "material density scales must be positive and finite");
    if (m.woodcock) {
        // transport/kernels.py:1850-1872 for every occurring material
        for (int k = 0; k < m.M; ++k) {
            if (m.max_density[k] <= 0.0) continue;
            for (int e = 0; e < m.K; ++e) {
                const double rate = density[k] * (t->absorption[k * m.K + e] + t->scattering[k * m.K + e]);
                if (!std::isfinite(rate) || rate >= t->majorant[e])
                    return tr_fail(DPT_E_INVALID, "material rate reaches the Woodcock majorant at an energy node");
            }
        }
    }
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(source->source[i]) || !std::isfinite(source->corner[i]) || !std::isfinite(source->u[i]) ||
            !std::isfinite(source->v[i]))
            return tr_fail(DPT_E_INVALID, "source geometry must be finite");
    Call c;
    std::memset(&c, 0, sizeof c);
    for (int k = 0; k < m.M; ++k) c.density[k] = density[k];
    std::memcpy(c.src, source->source, sizeof c.src);
    std::memcpy(c.corner, source->corner, sizeof c.corner);
    std::memcpy(c.su, source->u, sizeof c.su);
    std::memcpy(c.sv, source->v, sizeof c.sv);
    std::memcpy(c.extent, source->extent, sizeof c.extent);
    c.weight = source->weight;
    c.amplitude = amplitude;
    c.seed = seed;
    c.first = first_history;
    c.count = count;
    c.ids = t->d_ids;
    c.tables = t->d_tables;
    if (records != nullptr) {
        c.pixel = records->pixel;
        c.score = records->score;
        c.energy = records->energy;
        c.events = records->events;
        c.real_events = records->real_events;
        c.status = records->status;
    }
    c.tally = tally;
    c.histogram = t->d_histogram;
    c.failure_bits = t->d_bits;
    c.counter = t->d_counter;
    DeviceGuard guard(t->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t waves = (count + 63) / 64;
    const int grid = t->waves == 2 ? t->grid2 : t->grid4;
    const int blocks = (int)std::min<int64_t>(grid, (waves + 3) / 4);
    const size_t lds = (size_t)m.table_doubles * sizeof(double);
    zero_counter_kernel<<<1, 1, 0, s>>>(t->d_counter);
    if (t->waves == 2)
        trace_kernel<2><<<blocks, kBlock, lds, s>>>(m, c);
    else
        trace_kernel<4><<<blocks, kBlock, lds, s>>>(m, c);
    if (hipGetLastError() != hipSuccess) return tr_fail(DPT_E_DEVICE, "trace launch failed");
    return DPT_OK;
}

int dpt_transport_moments(dpt_transport *t, const double *tally, int64_t histories, double *mean, double *variance,
                          void *stream) {
    if (t == nullptr || tally == nullptr || mean == nullptr || variance == nullptr)
        return tr_fail(DPT_E_INVALID, "transport handle, tally and outputs are required");
    if (histories < 2) return tr_fail(DPT_E_INVALID, "moments need at least two histories");
    const int64_t pixels = (int64_t)t->model.det_w * t->model.det_h;
    DeviceGuard guard(t->device);
    moments_kernel<<<(unsigned)((pixels + 255) / 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
        tally, pixels, (double)histories, mean, variance);
    if (hipGetLastError() != hipSuccess) return tr_fail(DPT_E_DEVICE, "moments launch failed");
    return DPT_OK;
}

int dpt_transport_status(dpt_transport *t, int64_t *histogram, int32_t *failure_bits, int32_t clear, void *stream) {
    if (t == nullptr) return tr_fail(DPT_E_INVALID, "transport handle is required");
    DeviceGuard guard(t->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(t->h_histogram, t->d_histogram, 7 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(t->h_histogram + 7, t->d_bits, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
        return tr_fail(DPT_E_DEVICE, "status readback failed");
    if (hipStreamSynchronize(s) != hipSuccess) return tr_fail(DPT_E_DEVICE, "transport stream failed");
    if (histogram != nullptr)
        for (int k = 0; k < 7; ++k) histogram[k] = (int64_t)t->h_histogram[k];
    if (failure_bits != nullptr) {
        int bits;
        std::memcpy(&bits, &t->h_histogram[7], sizeof bits);
        *failure_bits = bits;
    }
    if (clear) {
        if (hipMemsetAsync(t->d_histogram, 0, 8 * sizeof(unsigned long long), s) != hipSuccess ||
            hipMemsetAsync(t->d_bits, 0, sizeof(int), s) != hipSuccess)
            return tr_fail(DPT_E_DEVICE, "status clear failed");
    }
    return DPT_OK;
}

}  // extern "C"
