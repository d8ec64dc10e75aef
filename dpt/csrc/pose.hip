// SE(3) chart on the device: G(xi) = anchor @ exp(hat(xi)) and its
// gradients, one launch each way (python/dpt/torch/geometry.py:19-63 forms
// the same 4x4 twist and takes torch.linalg.matrix_exp of it, whose many
// small launches and host round trips dominated a one-view request).
//
// Packed poses: 12 FP64, row-major rotation then translation (mm).
// xi = (tx, ty, tz, rx, ry, rz); hat(xi) = [[W, rho], [0, 0]] with
// W = [[0, -rz, ry], [rz, 0, -rx], [-ry, rx, 0]], rho = (tx, ty, tz).
//
// exp(M) by scaling and squaring: X = M / 2^s with ||X||_inf <= 1/2, the
// degree-18 Taylor polynomial in Horner form (remainder below
// 0.5^19 / 19! ~ 2e-23 relative), then s squarings.  The partial derivative
// along generator E_k = d hat / d xi_k is carried through the same steps
// (forward mode: d(AB) = dA B + A dB), so it is the exact derivative of the
// computed map.  One thread per (pose, k).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/dpt_amd.h"

namespace {

constexpr int kJet = 84;  // R (9), t (3), then d(R, t)/d xi_k (12) for k = 0..5

__device__ __forceinline__ void matmul4(const double *a, const double *b, double *c) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double s = a[4 * i] * b[j];
#pragma unroll
            for (int k = 1; k < 4; ++k) s = fma(a[4 * i + k], b[4 * k + j], s);
            c[4 * i + j] = s;
        }
}

__global__ __launch_bounds__(256) void compose_kernel(const double *__restrict__ anchor,
                                                      const double *__restrict__ tangent, int64_t count,
                                                      double *__restrict__ out, double *__restrict__ jet) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * 6) return;
    const int64_t i = t / 6;
    const int k = (int)(t - i * 6);
    const double *xi = tangent + 6 * i;
    double m[16], e[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) m[q] = e[q] = 0.0;
    m[1] = -xi[5];
    m[2] = xi[4];
    m[4] = xi[5];
    m[6] = -xi[3];
    m[8] = -xi[4];
    m[9] = xi[3];
    m[3] = xi[0];
    m[7] = xi[1];
    m[11] = xi[2];
    // generator of xi_k
    if (k < 3) {
        e[4 * k + 3] = 1.0;
    } else if (k == 3) {
        e[6] = -1.0;
        e[9] = 1.0;
    } else if (k == 4) {
        e[2] = 1.0;
        e[8] = -1.0;
    } else {
        e[1] = -1.0;
        e[4] = 1.0;
    }
    double norm = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) norm = fmax(norm, fabs(m[4 * r]) + fabs(m[4 * r + 1]) + fabs(m[4 * r + 2]) + fabs(m[4 * r + 3]));
    int s = 0;
    double scale = 1.0;
    while (norm > 0.5 && s < 1100) {  // NaN skips; powers of two scale exactly
        norm *= 0.5;
        scale *= 0.5;
        ++s;
    }
    double x[16], dx[16], p[16], dp[16], q[16], dq[16], w[16];
#pragma unroll
    for (int a = 0; a < 16; ++a) {
        x[a] = m[a] * scale;
        dx[a] = e[a] * scale;
        p[a] = (a % 5 == 0) ? 1.0 : 0.0;
        dp[a] = 0.0;
    }
    for (int j = 18; j >= 1; --j) {
        const double inv = 1.0 / (double)j;
        matmul4(x, p, q);
        matmul4(dx, p, dq);
        matmul4(x, dp, w);
#pragma unroll
        for (int a = 0; a < 16; ++a) {
            p[a] = ((a % 5 == 0) ? 1.0 : 0.0) + q[a] * inv;
            dp[a] = (dq[a] + w[a]) * inv;
        }
    }
    for (int r = 0; r < s; ++r) {
        matmul4(dp, p, dq);
        matmul4(p, dp, w);
        matmul4(p, p, q);
#pragma unroll
        for (int a = 0; a < 16; ++a) {
            dp[a] = dq[a] + w[a];
            p[a] = q[a];
        }
    }
    double *jt = jet + kJet * i;
    if (k == 0) {
        const double *ra = anchor + 12 * i;
        double *o = out + 12 * i;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                jt[3 * r + c] = p[4 * r + c];
                o[3 * r + c] = fma(ra[3 * r + 2], p[8 + c], fma(ra[3 * r + 1], p[4 + c], ra[3 * r] * p[c]));
            }
            jt[9 + r] = p[4 * r + 3];
            o[9 + r] = fma(ra[3 * r + 2], p[11], fma(ra[3 * r + 1], p[7], ra[3 * r] * p[3])) + ra[9 + r];
        }
    }
    double *d = jt + 12 + 12 * k;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) d[3 * r + c] = dp[4 * r + c];
        d[9 + r] = dp[4 * r + 3];
    }
}

// composed = Ra R, translation = Ra t + ta: dRa = G R^T + gt t^T, dta = gt,
// d xi_k = <Ra^T G, dR_k> + <Ra^T gt, dt_k>
__global__ __launch_bounds__(256) void compose_vjp_kernel(const double *__restrict__ anchor,
                                                          const double *__restrict__ jet,
                                                          const double *__restrict__ grad, int64_t count,
                                                          double *__restrict__ grad_anchor,
                                                          double *__restrict__ grad_tangent) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const double *ra = anchor + 12 * i, *jt = jet + kJet * i, *g = grad + 12 * i;
    double *ga = grad_anchor + 12 * i, *gx = grad_tangent + 6 * i;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            double v = g[9 + a] * jt[9 + b];
#pragma unroll
            for (int c = 0; c < 3; ++c) v = fma(g[3 * a + c], jt[3 * b + c], v);
            ga[3 * a + b] = v;
        }
        ga[9 + a] = g[9 + a];
    }
    double h[9], ht[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
#pragma unroll
        for (int c = 0; c < 3; ++c) h[3 * b + c] = ra[b] * g[c] + ra[3 + b] * g[3 + c] + ra[6 + b] * g[6 + c];
        ht[b] = ra[b] * g[9] + ra[3 + b] * g[10] + ra[6 + b] * g[11];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double *d = jt + 12 + 12 * k;
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) v = fma(h[q], d[q], v);
#pragma unroll
        for (int q = 0; q < 3; ++q) v = fma(ht[q], d[9 + q], v);
        gx[k] = v;
    }
}

}  // namespace

extern "C" int dpt_compose_pose(const double *anchor, const double *tangent, int64_t count, double *out, double *jet,
                                void *stream) {
    if (count < 0 || (count > 0 && (anchor == nullptr || tangent == nullptr || out == nullptr || jet == nullptr)))
        return DPT_E_INVALID;
    if (count == 0) return DPT_OK;
    const int64_t threads = count * 6;
    hipLaunchKernelGGL(compose_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), anchor, tangent, count, out, jet);
    return hipGetLastError() == hipSuccess ? DPT_OK : DPT_E_DEVICE;
}

extern "C" int dpt_compose_pose_vjp(const double *anchor, const double *jet, const double *grad_out, int64_t count,
                                    double *grad_anchor, double *grad_tangent, void *stream) {
    if (count < 0 || (count > 0 && (anchor == nullptr || jet == nullptr || grad_out == nullptr ||
                                    grad_anchor == nullptr || grad_tangent == nullptr)))
        return DPT_E_INVALID;
    if (count == 0) return DPT_OK;
    hipLaunchKernelGGL(compose_vjp_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), anchor, jet, grad_out, count, grad_anchor, grad_tangent);
    return hipGetLastError() == hipSuccess ? DPT_OK : DPT_E_DEVICE;
}
