// probe.hip — libprobe.so: a vector-issue probe for the benchmark's rooflines
// (not part of the engine).  k_valu_probe keeps eight independent FMA chains
// per lane, enough waves on every SIMD to hide the FMA latency, so it issues one
// VALU instruction per SIMD slot; its SQ_INSTS_VALU x 64 / launch time is the
// sustained VALU issue rate of the device at its running clock.
#include <hip/hip_runtime.h>

#include <cstdint>

__global__ __launch_bounds__(256) void k_valu_probe(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float a0 = gid * 1e-7f, a1 = a0 + 1.f, a2 = a0 + 2.f, a3 = a0 + 3.f;
    float a4 = a0 + 4.f, a5 = a0 + 5.f, a6 = a0 + 6.f, a7 = a0 + 7.f;
    for (int i = 0; i < iters; ++i) {
        a0 = __builtin_fmaf(a0, m, c); a1 = __builtin_fmaf(a1, m, c);
        a2 = __builtin_fmaf(a2, m, c); a3 = __builtin_fmaf(a3, m, c);
        a4 = __builtin_fmaf(a4, m, c); a5 = __builtin_fmaf(a5, m, c);
        a6 = __builtin_fmaf(a6, m, c); a7 = __builtin_fmaf(a7, m, c);
    }
    out[gid] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}

// One launch of `blocks` x 256 threads running `iters` rounds of 8 FMAs; the
// launch is timed with HIP events (after one untimed warm launch).  Returns 0,
// or the HIP error code.
extern "C" int rzk_probe_valu(int blocks, int iters, float* out_ms) {
    if (blocks <= 0 || iters <= 0 || !out_ms) return (int)hipErrorInvalidValue;
    float* d = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipMalloc((void**)&d, sizeof(float) * (size_t)blocks * 256);
    if (err == hipSuccess) err = hipStreamCreate(&st);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, st, d, iters, 0.999f, 1e-3f);
        (void)hipEventRecord(e0, st);
        hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, st, d, iters, 0.999f, 1e-3f);
        (void)hipEventRecord(e1, st);
        err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipGetLastError();
        if (err == hipSuccess) err = hipEventElapsedTime(out_ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (d) (void)hipFree(d);
    return (int)err;
}
