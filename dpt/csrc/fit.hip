// Batched pose-fit solves on resident volumes (the serving payload).
//
// One step advances every active request by one Levenberg-Marquardt iteration
// of   min_xi  1/2 sum_rays (L_ray(pose o exp(xi^)) - target_ray)^2,
// where L is the reference's FP32 clipped-midpoint optical depth
// (kernels/projection_fp32.py:334-374) and the Jacobian row of each ray is the
// reference's local-right pose derivative with a unit seed
// (kernels/projection_fp32.py:549-602, kernels/geometry.py:25-52).  No
// reference counterpart exists for the batching itself; the normal-equation
// solve follows the Gauss-Newton pose step of python/dpt/pose_optimization.py.
//
// fit_step_kernel: persistent workgroups pull 16x16 detector tiles of any
//   request from a device tile table (dynamic counter).  Per ray ONE march
//   yields both L (so the residual) and the seed-independent pose accumulators
//   (mean, sum grad, sum t*grad, clip-face shifts); FP32 sampling with
//   compensated along-ray sums, FP32 or BF16/FP16 storage.  Each wave sums
//   J^T J (21), J^T r (6) and r^2 in FP64 and writes one partial.
// fit_update_kernel: one workgroup per request sums its partials in a fixed
//   order (results do not depend on the batch), then one lane solves the 6x6
//   damped normal equations, accepts or rejects the last step and writes the
//   next pose.  Grid sizes and table contents are read from device memory, so
//   a step can be captured once per grid bucket as a hipGraph and replayed.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "device_common.h"

namespace dpt {

constexpr int kFitTerms = 28;         // 21 (J^T J upper) + 6 (J^T r) + 1 (r^2)
constexpr int kMaxFitRequests = 256;

struct FitRequestDev {
    Params P;                          // grid + detector of this request (panels, views, rows, cols, samples)
    const void *field;
    int storage;
    int tile_begin;
    const float *targets;
    double *state;
};

struct FitTable {
    int n_requests;
    int total_tiles;
    int pad0, pad1;
    int tile_begin[kMaxFitRequests + 1];
    FitRequestDev req[kMaxFitRequests];
};

// Ray march of one request ray: L and the 6-component unit-seed Jacobian row.
// FP32 storage sums plainly (the projector's FP32 VJP arithmetic); 16-bit
// storage uses blocked compensated sums.
template <typename T, bool COMPENSATED>
__device__ __forceinline__ void fit_ray(const Params &P, const T *field, bool cells, const double *pose, int view, int row,
                                        int col, float &L, double *j, bool &valid) {
    World w;
    ray_world(P, pose, view, row, col, w);
    Ray r;
    ray_grid(P, w, r);
    float o[3] = {(float)r.o[0], (float)r.o[1], (float)r.o[2]};
    float d[3] = {(float)r.d[0], (float)r.d[1], (float)r.d[2]};
    const float len = (float)r.len;
    L = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; ++i) j[i] = 0.0;
    valid = r.valid && finite3f(o[0], o[1], o[2]) && finite3f(d[0], d[1], d[2]) && isfinite(len) && len > 0.0f;
    if (!valid) return;
    const Interval32 iv = interval32(P, o, d);
    if (!(iv.upper > iv.lower)) return;
    const float span = iv.upper - iv.lower;
    const float step = span / (float)P.samples;
    Acc<COMPENSATED> mean, dor0, dor1, dor2, shl, shu, dd0, dd1, dd2;
    mean.init();
    dor0.init();
    dor1.init();
    dor2.init();
    shl.init();
    shu.init();
    dd0.init();
    dd1.init();
    dd2.init();
    float v[8];
    int cur = -1;
    // per-sample support tests only for waves with a ray whose end samples
    // leave the support (uniform flag)
    const bool checked = __builtin_amdgcn_ballot_w64(!span_inside(P, o, d, iv.lower, step, 0, P.samples)) != 0;
    for (int k = 0; k < P.samples; ++k) {
        if constexpr (COMPENSATED) {
            if (k > 0 && (k % kFoldBlock) == 0) {
                mean.fold();
                dor0.fold();
                dor1.fold();
                dor2.fold();
                shl.fold();
                shu.fold();
                dd0.fold();
                dd1.fold();
                dd2.fold();
            }
        }
        float pos;
        Cell32 c;
        if (checked) {
            sample_at<true>(P, o, d, iv.lower, step, k, pos, c);
            if (c.base < 0) continue;
        } else {
            sample_at<false>(P, o, d, iv.lower, step, k, pos, c);
        }
        if (c.base != cur) {
            if (cells)
                load8<T, kLayoutCells>(P, field, c.base, v);
            else if (P.sx)
                load8<T, kLayoutPair>(P, field, c.base, v);
            else
                load8<T, kLayoutScalar>(P, field, c.base, v);
            cur = c.base;
        }
        float wt[8], g[3];
        weights32(c, wt);
        mean.add(interpolate32(wt, v));
        slopes32(c, v, g);
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
    }
    const float m = mean.value();
    L = m * (len * step);
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
    pose_epilogue<6>(P, pose, w, (double)len / (double)P.samples, og, dg, j);
}

__global__ __launch_bounds__(kBlock) void fit_step_kernel(const FitTable *__restrict__ tab,
                                                          double *__restrict__ partials, int *counter) {
    __shared__ int next;
    const int total = tab->total_tiles;
    const int n = tab->n_requests;
    while (true) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int t = next;
        __syncthreads();
        if (t >= total) break;
        // request owning tile t: largest r with tile_begin[r] <= t
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tab->tile_begin[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const FitRequestDev &R = tab->req[lo];
        const Params &P = R.P;
        const int local = t - tab->tile_begin[lo];
        const int view = local / P.tiles_per_group;
        const int rem = local - view * P.tiles_per_group;
        const int ty = rem / P.tiles_x, tx = rem - (rem / P.tiles_x) * P.tiles_x;
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
        const int row = ty * kTile + (w >> 1) * 8 + (l >> 3);
        const int col = tx * kTile + (w & 1) * 8 + (l & 7);
        double acc[kFitTerms];
#pragma unroll
        for (int i = 0; i < kFitTerms; ++i) acc[i] = 0.0;
        if (row < P.rows && col < P.cols) {
            float L;
            double jr[6];
            bool valid;
            const int element = R.storage & ~DPT_STORE_CELLS;
            const bool cells = (R.storage & DPT_STORE_CELLS) != 0;
            if (element == DPT_STORE_BF16)
                fit_ray<__hip_bfloat16, true>(P, static_cast<const __hip_bfloat16 *>(R.field), cells, R.state, view,
                                              row, col, L, jr, valid);
            else if (element == DPT_STORE_F16)
                fit_ray<__half, true>(P, static_cast<const __half *>(R.field), cells, R.state, view, row, col, L, jr,
                                      valid);
            else
                fit_ray<float, false>(P, static_cast<const float *>(R.field), cells, R.state, view, row, col, L, jr,
                                      valid);
            const double res = (double)L - (double)R.targets[((size_t)view * P.rows + row) * P.cols + col];
            int q = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a)
#pragma unroll
                for (int b = a; b < 6; ++b) acc[q++] = jr[a] * jr[b];
#pragma unroll
            for (int a = 0; a < 6; ++a) acc[21 + a] = jr[a] * res;
            acc[27] = res * res;
        }
        double *out = partials + ((size_t)t * 4 + w) * kFitTerms;
#pragma unroll
        for (int i = 0; i < kFitTerms; ++i) {
            const double s = wave_sum(acc[i]);
            if (l == 0) out[i] = s;
        }
    }
}

// SE(3) exponential of xi = (rho, omega), 12 packed values (row-major R, t).
__device__ void se3_exp(const double *xi, double *out) {
    const double wx = xi[3], wy = xi[4], wz = xi[5];
    const double t2 = wx * wx + wy * wy + wz * wz;
    double a, b, c;
    if (t2 < 1e-2) {
        a = 1 - t2 / 6 + t2 * t2 / 120 - t2 * t2 * t2 / 5040;
        b = 0.5 - t2 / 24 + t2 * t2 / 720 - t2 * t2 * t2 / 40320;
        c = 1.0 / 6 - t2 / 120 + t2 * t2 / 5040 - t2 * t2 * t2 / 362880;
    } else {
        const double t = sqrt(t2);
        a = sin(t) / t;
        b = (1 - cos(t)) / t2;
        c = (t - sin(t)) / (t2 * t);
    }
    const double W[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
    double W2[9];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) W2[3 * i + k] = W[3 * i] * W[k] + W[3 * i + 1] * W[3 + k] + W[3 * i + 2] * W[6 + k];
    double V[9];
    for (int i = 0; i < 9; ++i) {
        const double id = (i % 4 == 0) ? 1.0 : 0.0;
        out[i] = id + a * W[i] + b * W2[i];
        V[i] = id + b * W[i] + c * W2[i];
    }
    for (int i = 0; i < 3; ++i) out[9 + i] = V[3 * i] * xi[0] + V[3 * i + 1] * xi[1] + V[3 * i + 2] * xi[2];
}

// Solve (6x6 SPD) M x = y by Cholesky; returns false when M is not SPD.
__device__ bool solve6(double *M, const double *y, double *x) {
    for (int i = 0; i < 6; ++i) {
        for (int k = 0; k < i; ++k) {
            double s = M[6 * i + k];
            for (int m = 0; m < k; ++m) s -= M[6 * i + m] * M[6 * k + m];
            M[6 * i + k] = s / M[6 * k + k];
        }
        double s = M[6 * i + i];
        for (int m = 0; m < i; ++m) s -= M[6 * i + m] * M[6 * i + m];
        if (!(s > 0.0)) return false;
        M[6 * i + i] = sqrt(s);
    }
    double z[6];
    for (int i = 0; i < 6; ++i) {
        double s = y[i];
        for (int m = 0; m < i; ++m) s -= M[6 * i + m] * z[m];
        z[i] = s / M[6 * i + i];
    }
    for (int i = 5; i >= 0; --i) {
        double s = z[i];
        for (int m = i + 1; m < 6; ++m) s -= M[6 * m + i] * x[m];
        x[i] = s / M[6 * i + i];
    }
    return true;
}

__global__ __launch_bounds__(kBlock) void fit_update_kernel(const FitTable *__restrict__ tab,
                                                            const double *__restrict__ partials, int *counter) {
    const int r = blockIdx.x;
    if (r == 0 && threadIdx.x == 0) *counter = 0;  // re-arm the tile counter for the next step
    if (r >= tab->n_requests) return;
    const size_t begin = (size_t)tab->tile_begin[r] * 4, end = (size_t)tab->tile_begin[r + 1] * 4;
    __shared__ double red[kBlock];
    __shared__ double sums[kFitTerms];
    for (int i = 0; i < kFitTerms; ++i) {
        double v = 0.0;
        for (size_t p = begin + threadIdx.x; p < end; p += kBlock) v += partials[p * kFitTerms + i];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) sums[i] = red[0];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double *st = tab->req[r].state;
    const double loss = 0.5 * sums[27];
    double A[21], b[6];
    for (int i = 0; i < 21; ++i) A[i] = sums[i];
    for (int i = 0; i < 6; ++i) b[i] = sums[21 + i];
    const double evals = st[DPT_FIT_EVALS];
    double lambda = st[DPT_FIT_LAMBDA];
    if (evals == 0.0 || (isfinite(loss) && loss <= st[DPT_FIT_BEST_LOSS])) {
        for (int i = 0; i < 12; ++i) st[DPT_FIT_BEST_POSE + i] = st[DPT_FIT_POSE + i];
        st[DPT_FIT_BEST_LOSS] = loss;
        for (int i = 0; i < 21; ++i) st[DPT_FIT_JTJ + i] = A[i];
        for (int i = 0; i < 6; ++i) st[DPT_FIT_JTR + i] = b[i];
        st[DPT_FIT_ACCEPTED] += 1.0;
        lambda = evals == 0.0 ? 1e-3 : fmax(lambda * 0.1, 1e-12);
    } else {
        lambda = fmin(lambda * 10.0, 1e12);
        for (int i = 0; i < 21; ++i) A[i] = st[DPT_FIT_JTJ + i];
        for (int i = 0; i < 6; ++i) b[i] = st[DPT_FIT_JTR + i];
    }
    st[DPT_FIT_LAST_LOSS] = loss;
    st[DPT_FIT_EVALS] = evals + 1.0;
    st[DPT_FIT_LAMBDA] = lambda;
    // damped normal equations from the best pose
    double M[36], y[6], delta[6] = {0, 0, 0, 0, 0, 0};
    int q = 0;
    for (int a = 0; a < 6; ++a)
        for (int c = a; c < 6; ++c) {
            M[6 * a + c] = A[q];
            M[6 * c + a] = A[q];
            ++q;
        }
    double trace = 0.0;
    for (int a = 0; a < 6; ++a) trace += M[7 * a];
    for (int a = 0; a < 6; ++a) {
        M[7 * a] += lambda * M[7 * a] + 1e-12 * trace + 1e-300;
        y[a] = -b[a];
    }
    if (!solve6(M, y, delta))
        for (int a = 0; a < 6; ++a) delta[a] = 0.0;
    double inc[12];
    se3_exp(delta, inc);
    const double *G = st + DPT_FIT_BEST_POSE;
    double next[12];
    for (int i = 0; i < 3; ++i) {
        for (int c = 0; c < 3; ++c)
            next[3 * i + c] = G[3 * i] * inc[c] + G[3 * i + 1] * inc[3 + c] + G[3 * i + 2] * inc[6 + c];
        next[9 + i] = G[3 * i] * inc[9] + G[3 * i + 1] * inc[10] + G[3 * i + 2] * inc[11] + G[9 + i];
    }
    for (int i = 0; i < 12; ++i) st[DPT_FIT_POSE + i] = next[i];
}

}  // namespace dpt

// ================================================================ C ABI ====

using namespace dpt;

struct dpt_fit {
    int device = 0;
    int max_requests = 0;
    int64_t max_tiles = 0;
    FitTable *d_table = nullptr;
    FitTable *h_table[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {nullptr, nullptr};
    int flip = 0;
    double *d_partials = nullptr;
    int *d_counter = nullptr;
    std::map<std::pair<int, hipStream_t>, hipGraphExec_t> graphs;
    std::map<int, float *> alpha;  // sample-fraction tables by samples per ray
};

namespace {

thread_local std::string g_fit_error;

int fit_fail(int code, const std::string &msg) {
    g_fit_error = msg;
    return code;
}

// Grid geometry of one request (same rules as dpt_plan_create).
bool storage_ok(int storage) {
    const int element = storage & ~DPT_STORE_CELLS;
    return element == DPT_STORE_F32 || element == DPT_STORE_BF16 || element == DPT_STORE_F16;
}

bool fill_params(Params &P, const dpt_volume_desc &v, const dpt_fit_request &q, std::string &why) {
    if (v.nx < 1 || v.ny < 1 || v.nz < 1 || (int64_t)v.nx * v.ny * v.nz > INT32_MAX) {
        why = "volume counts must be positive and fit signed 32-bit indexing";
        return false;
    }
    if ((long long)v.ny * v.nz > kMaxSliceProduct) {
        why = "volume ny * nz exceeds 2^24 (24-bit cell index arithmetic)";
        return false;
    }
    if (q.views < 1 || q.rows < 1 || q.cols < 1 || q.samples < 1) {
        why = "request views, rows, cols and samples must be positive";
        return false;
    }
    std::memset(&P, 0, sizeof P);
    P.nx = v.nx;
    P.ny = v.ny;
    P.nz = v.nz;
    // gather index: voxels of the linear field, or cells of the cell layout
    const bool cells = (v.storage & DPT_STORE_CELLS) != 0;
    P.ix = cells ? (v.nx > 1 ? v.nx - 1 : 1) : v.nx;
    P.iy = cells ? (v.ny > 1 ? v.ny - 1 : 1) : v.ny;
    P.sx = v.nx > 1 ? 1 : 0;
    P.sy = v.ny > 1 ? v.nx : 0;
    P.sz = v.nz > 1 ? v.nx * v.ny : 0;
    P.hx = (float)v.nx - 0.5f;
    P.hy = (float)v.ny - 0.5f;
    P.hz = (float)v.nz - 0.5f;
    P.cx = (float)(v.nx - 1);
    P.cy = (float)(v.ny - 1);
    P.cz = (float)(v.nz - 1);
    P.capx = v.nx - 2 > 0 ? v.nx - 2 : 0;
    P.capy = v.ny - 2 > 0 ? v.ny - 2 : 0;
    P.capz = v.nz - 2 > 0 ? v.nz - 2 : 0;
    std::memcpy(P.g, v.object_to_grid, sizeof P.g);
    std::memcpy(P.go, v.grid_origin, sizeof P.go);
    P.views = q.views;
    P.rows = q.rows;
    P.cols = q.cols;
    P.samples = q.samples;
    P.vb = 1;
    P.tile_w = P.tile_h = kTile;
    P.tiles_x = (q.cols + kTile - 1) / kTile;
    P.tiles_y = (q.rows + kTile - 1) / kTile;
    P.tiles_per_group = P.tiles_x * P.tiles_y;
    P.panels = q.panels;
    return true;
}

}  // namespace

extern "C" {

const char *dpt_fit_last_error(void) { return g_fit_error.c_str(); }

int dpt_fit_destroy(dpt_fit *f) {
    if (f == nullptr) return DPT_OK;
    int previous = 0;
    (void)hipGetDevice(&previous);
    (void)hipSetDevice(f->device);
    for (auto &kv : f->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto &kv : f->alpha) (void)hipFree(kv.second);
    (void)hipFree(f->d_table);
    (void)hipFree(f->d_partials);
    (void)hipFree(f->d_counter);
    for (int i = 0; i < 2; ++i) {
        if (f->h_table[i]) (void)hipHostFree(f->h_table[i]);
        if (f->copied[i]) (void)hipEventDestroy(f->copied[i]);
    }
    (void)hipSetDevice(previous);
    delete f;
    return DPT_OK;
}

int dpt_fit_create(int32_t device, int32_t max_requests, int64_t max_tiles, dpt_fit **out) {
    if (out == nullptr) return fit_fail(DPT_E_INVALID, "out is required");
    *out = nullptr;
    if (max_requests < 1 || max_requests > kMaxFitRequests)
        return fit_fail(DPT_E_INVALID, "max_requests must be in [1, 256]");
    if (max_tiles < 1) return fit_fail(DPT_E_INVALID, "max_tiles must be positive");
    int previous = 0;
    (void)hipGetDevice(&previous);
    if (hipSetDevice(device) != hipSuccess) return fit_fail(DPT_E_DEVICE, "hipSetDevice failed");
    dpt_fit *f = new dpt_fit();
    f->device = device;
    f->max_requests = max_requests;
    f->max_tiles = max_tiles;
    bool ok = hipMalloc((void **)&f->d_table, sizeof(FitTable)) == hipSuccess &&
              hipMalloc((void **)&f->d_partials, sizeof(double) * kFitTerms * 4 * (size_t)max_tiles) == hipSuccess &&
              hipMalloc((void **)&f->d_counter, sizeof(int)) == hipSuccess &&
              hipMemset(f->d_counter, 0, sizeof(int)) == hipSuccess;
    for (int i = 0; i < 2 && ok; ++i)
        ok = hipHostMalloc((void **)&f->h_table[i], sizeof(FitTable), hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&f->copied[i], hipEventDisableTiming) == hipSuccess;
    (void)hipSetDevice(previous);
    if (!ok) {
        dpt_fit_destroy(f);
        return fit_fail(DPT_E_MEMORY, "allocation of the fit tables failed");
    }
    *out = f;
    return DPT_OK;
}

int dpt_fit_step(dpt_fit *f, const dpt_volume_desc *volumes, int32_t n_volumes, const dpt_fit_request *requests,
                 int32_t n_requests, int32_t use_graph, void *stream) {
    if (f == nullptr || (n_requests > 0 && (requests == nullptr || volumes == nullptr)))
        return fit_fail(DPT_E_INVALID, "fit, volumes and requests are required");
    if (n_requests < 0 || n_requests > f->max_requests)
        return fit_fail(DPT_E_INVALID, "n_requests exceeds the fit capacity");
    if (n_requests == 0) return DPT_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int previous = 0;
    (void)hipGetDevice(&previous);
    (void)hipSetDevice(f->device);
    // double-buffered pinned staging: wait until this buffer's last copy is done
    const int buf = f->flip;
    f->flip ^= 1;
    (void)hipEventSynchronize(f->copied[buf]);
    FitTable *h = f->h_table[buf];
    int64_t tiles = 0;
    std::string why;
    for (int r = 0; r < n_requests; ++r) {
        const dpt_fit_request &q = requests[r];
        if (q.volume < 0 || q.volume >= n_volumes || q.panels == nullptr || q.targets == nullptr ||
            q.state == nullptr) {
            (void)hipSetDevice(previous);
            return fit_fail(DPT_E_INVALID, "request " + std::to_string(r) + ": invalid volume index or pointer");
        }
        const dpt_volume_desc &v = volumes[q.volume];
        FitRequestDev &R = h->req[r];
        if (!fill_params(R.P, v, q, why) || v.field == nullptr ||
            !storage_ok(v.storage)) {
            (void)hipSetDevice(previous);
            return fit_fail(DPT_E_INVALID, "request " + std::to_string(r) + ": " + (why.empty() ? "invalid volume" : why));
        }
        auto table = f->alpha.find(q.samples);
        if (table == f->alpha.end()) {
            std::vector<float> host(sample_table_entries(q.samples));
            fill_alpha_table(host.data(), q.samples);
            float *dev = nullptr;
            if (hipMalloc(&dev, sizeof(float) * host.size()) != hipSuccess ||
                hipMemcpy(dev, host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipFree(dev);
                (void)hipSetDevice(previous);
                return fit_fail(DPT_E_MEMORY, "sample-fraction table allocation failed");
            }
            table = f->alpha.emplace(q.samples, dev).first;
        }
        R.P.alpha = table->second;
        R.field = v.field;
        R.storage = v.storage;
        R.targets = q.targets;
        R.state = q.state;
        R.tile_begin = (int)tiles;
        h->tile_begin[r] = (int)tiles;
        tiles += (int64_t)R.P.tiles_per_group * q.views;
    }
    if (tiles > f->max_tiles || tiles > INT32_MAX) {
        (void)hipSetDevice(previous);
        return fit_fail(DPT_E_INVALID, "active requests exceed the fit tile capacity");
    }
    h->tile_begin[n_requests] = (int)tiles;
    h->n_requests = n_requests;
    h->total_tiles = (int)tiles;
    const size_t bytes = offsetof(FitTable, req) + sizeof(FitRequestDev) * n_requests;
    hipError_t e = hipMemcpyAsync(f->d_table, h, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(f->copied[buf], s);
    if (e != hipSuccess) {
        (void)hipSetDevice(previous);
        return fit_fail(DPT_E_DEVICE, std::string("table upload: ") + hipGetErrorString(e));
    }
    // persistent grid, bucketed by work so small batches launch few workgroups
    int grid = 64;
    while (grid < tiles && grid < 2048) grid <<= 1;
    const int update_grid = f->max_requests;
    if (use_graph) {
        auto key = std::make_pair(grid, s);
        auto it = f->graphs.find(key);
        if (it == f->graphs.end()) {
            hipGraph_t graph;
            hipGraphExec_t exec;
            e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(fit_step_kernel, dim3(grid), dim3(kBlock), 0, s, f->d_table, f->d_partials,
                                   f->d_counter);
                hipLaunchKernelGGL(fit_update_kernel, dim3(update_grid), dim3(kBlock), 0, s, f->d_table, f->d_partials,
                                   f->d_counter);
                e = hipStreamEndCapture(s, &graph);
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e == hipSuccess) {
                (void)hipGraphDestroy(graph);
                it = f->graphs.emplace(key, exec).first;
            } else {
                (void)hipSetDevice(previous);
                return fit_fail(DPT_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
            }
        }
        e = hipGraphLaunch(it->second, s);
    } else {
        hipLaunchKernelGGL(fit_step_kernel, dim3(grid), dim3(kBlock), 0, s, f->d_table, f->d_partials, f->d_counter);
        hipLaunchKernelGGL(fit_update_kernel, dim3(update_grid), dim3(kBlock), 0, s, f->d_table, f->d_partials,
                           f->d_counter);
        e = hipGetLastError();
    }
    (void)hipSetDevice(previous);
    if (e != hipSuccess) return fit_fail(DPT_E_DEVICE, std::string("fit launch: ") + hipGetErrorString(e));
    return DPT_OK;
}

}  // extern "C"
