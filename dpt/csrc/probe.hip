// Issue-rate probe for the roofline report: the VALU ceiling against which
// the projector kernels' measured instruction rates are set.
//
// Each lane runs 16 independent dependency chains of one 32-bit VALU opcode
// (fused multiply-add, integer add) per loop iteration; 8 waves per SIMD are
// resident (256-lane workgroups, 2048 workgroups per 256 CUs), so issue, not
// latency, bounds the rate.  The loop body is inline assembly: exactly those
// instructions, no packed FP32 forms.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/dpt_amd.h"

namespace {

#define DPT_REP16(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15)

template <int KIND>
__global__ __launch_bounds__(256) void issue_probe_kernel(float *out, int iterations) {
    float r[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) r[i] = (float)threadIdx.x * 0.001f + (float)i;
    const float a = 1.0000001f, b = 1e-7f;
    for (int k = 0; k < iterations; ++k) {
#define DPT_FMA(i) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(r[i]) : "v"(a), "v"(b));
#define DPT_ADD(i) asm volatile("v_add_u32 %0, %0, %1" : "+v"(r[i]) : "v"(a));
        if constexpr (KIND == 0) {
            DPT_REP16(DPT_FMA)
        } else {
            DPT_REP16(DPT_ADD)
        }
#undef DPT_FMA
#undef DPT_ADD
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) s += r[i];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

extern "C" int dpt_valu_probe(int32_t kind, void *scratch_out, double *rate, void *stream) {
    if ((kind != 0 && kind != 1) || scratch_out == nullptr || rate == nullptr) return DPT_E_INVALID;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int blocks = 2048, iterations = 4000;
    hipEvent_t start, stop;
    if (hipEventCreate(&start) != hipSuccess) return DPT_E_DEVICE;
    if (hipEventCreate(&stop) != hipSuccess) {
        (void)hipEventDestroy(start);
        return DPT_E_DEVICE;
    }
    float *out = static_cast<float *>(scratch_out);
    int rc = DPT_OK;
    float ms = 0.0f;
    // one warm launch, then the timed one
    for (int pass = 0; pass < 2 && rc == DPT_OK; ++pass) {
        if (pass == 1 && hipEventRecord(start, s) != hipSuccess) rc = DPT_E_DEVICE;
        if (kind == 0)
            hipLaunchKernelGGL(issue_probe_kernel<0>, dim3(blocks), dim3(256), 0, s, out, iterations);
        else
            hipLaunchKernelGGL(issue_probe_kernel<1>, dim3(blocks), dim3(256), 0, s, out, iterations);
        if (hipGetLastError() != hipSuccess) rc = DPT_E_DEVICE;
    }
    if (rc == DPT_OK && (hipEventRecord(stop, s) != hipSuccess || hipEventSynchronize(stop) != hipSuccess ||
                         hipEventElapsedTime(&ms, start, stop) != hipSuccess || ms <= 0.0f))
        rc = DPT_E_DEVICE;
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    if (rc == DPT_OK) *rate = (double)blocks * 4 * iterations * 16 / (ms * 1e-3);
    return rc;
}
