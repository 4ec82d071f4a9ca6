// Measured instruction-throughput ceilings of the device, the denominators of the opacity
// build's compute roofline: the line loop spends one exponential per (line, cell), so its
// ceiling is the exponential rate, not HBM bandwidth.
//
//   k_exp32  raw v_exp_f32 (2^x) throughput: 8 independent chains x = 2^-x per lane, which
//            stay in [0.5, 1] and never touch denormals
//   k_exp64  throughput of the FP64 exp() instruction sequence the FP64 build uses
//
// The chains are independent so latency is hidden; every result is written out so no
// chain is dead code.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/slgpu.h"

namespace {

constexpr int P_BLOCKS = 4096, P_THREADS = 256, P_ILP = 8;
constexpr int P_ITER32 = 2048, P_ITER64 = 128;

__global__ __launch_bounds__(256) void k_exp32(float *out, float seed)
{
    float x[P_ILP];
    #pragma unroll
    for (int k = 0; k < P_ILP; ++k) x[k] = seed + 0.001f * (float)(k + (threadIdx.x & 7));
    for (int it = 0; it < P_ITER32; ++it) {
        #pragma unroll
        for (int k = 0; k < P_ILP; ++k) x[k] = __builtin_amdgcn_exp2f(-x[k]);
    }
    float s = 0.0f;
    #pragma unroll
    for (int k = 0; k < P_ILP; ++k) s += x[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_exp64(double *out, double seed)
{
    double x[P_ILP];
    #pragma unroll
    for (int k = 0; k < P_ILP; ++k) x[k] = seed + 0.001 * (double)(k + (threadIdx.x & 7));
    for (int it = 0; it < P_ITER64; ++it) {
        #pragma unroll
        for (int k = 0; k < P_ILP; ++k) x[k] = exp(-x[k]);
    }
    double s = 0.0;
    #pragma unroll
    for (int k = 0; k < P_ILP; ++k) s += x[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" int slgpu_peak_rates(double *out_exp32_per_s, double *out_exp64_per_s)
{
    const size_t n = (size_t)P_BLOCKS * P_THREADS;
    void *buf = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipMalloc(&buf, n * sizeof(double)) != hipSuccess) return SLGPU_ENOMEM;
    int rc = SLGPU_OK;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        rc = SLGPU_EHIP;
    }
    float best32 = 1e30f, best64 = 1e30f;
    for (int rep = 0; rep < 4 && rc == SLGPU_OK; ++rep) {
        float ms = 0.f;
        hipEventRecord(a, s);
        hipLaunchKernelGGL(k_exp32, dim3(P_BLOCKS), dim3(P_THREADS), 0, s, (float *)buf, 0.5f);
        hipEventRecord(b, s);
        if (hipEventSynchronize(b) != hipSuccess) { rc = SLGPU_EHIP; break; }
        hipEventElapsedTime(&ms, a, b);
        if (rep > 0 && ms < best32) best32 = ms;
        hipEventRecord(a, s);
        hipLaunchKernelGGL(k_exp64, dim3(P_BLOCKS), dim3(P_THREADS), 0, s, (double *)buf, 0.5);
        hipEventRecord(b, s);
        if (hipEventSynchronize(b) != hipSuccess) { rc = SLGPU_EHIP; break; }
        hipEventElapsedTime(&ms, a, b);
        if (rep > 0 && ms < best64) best64 = ms;
    }
    if (rc == SLGPU_OK && hipGetLastError() != hipSuccess) rc = SLGPU_EHIP;
    if (rc == SLGPU_OK) {
        if (out_exp32_per_s) *out_exp32_per_s = (double)n * P_ITER32 * P_ILP / (best32 * 1e-3);
        if (out_exp64_per_s) *out_exp64_per_s = (double)n * P_ITER64 * P_ILP / (best64 * 1e-3);
    }
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (s) hipStreamDestroy(s);
    hipFree(buf);
    return rc;
}
