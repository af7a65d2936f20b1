// probe.hip — libprobe.so: a vector-issue probe for the benchmark's rooflines
// (not part of the engine).  k_valu_probe keeps sixteen independent FMA chains
// per lane, enough waves on every SIMD to hide the FMA latency, so it issues one
// VALU instruction per SIMD slot; its SQ_INSTS_VALU x 64 / launch time is the
// sustained VALU issue rate of the device at its running clock.
#include <hip/hip_runtime.h>

#include <cstdint>

__global__ __launch_bounds__(256) void k_valu_probe(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = gid * 1e-7f + (float)k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) a[k] = __builtin_fmaf(a[k], m, c);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += a[k];
    out[gid] = s;
}

// The slab-test mix of a BVH box test: subtract, multiply, min, max, compare
// and select, sixteen independent chains per lane kept in [lo, hi].
__global__ __launch_bounds__(256) void k_valu_probe_slab(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = gid * 1e-7f + 0.01f * (float)k;
    const float lo = -4.f, hi = 4.f, y = 2.5f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            float t = (a[k] - c) * m;
            t = fmaxf(fminf(t, hi), lo);
            a[k] = t < y ? t : -t;
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += a[k];
    out[gid] = s;
}

// One launch of probe `kind` (0 FMA chains, 1 slab mix) over `blocks` x 256
// threads running `iters` rounds; the
// launch is timed with HIP events (after one untimed warm launch).  Returns 0,
// or the HIP error code.
extern "C" int rzk_probe_valu(int kind, int blocks, int iters, float* out_ms) {
    if (blocks <= 0 || iters <= 0 || !out_ms || kind < 0 || kind > 1) return (int)hipErrorInvalidValue;
    auto kern = kind == 0 ? k_valu_probe : k_valu_probe_slab;
    float* d = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipMalloc((void**)&d, sizeof(float) * (size_t)blocks * 256);
    if (err == hipSuccess) err = hipStreamCreate(&st);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, st, d, iters, 0.999f, 1e-3f);
        (void)hipEventRecord(e0, st);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, st, d, iters, 0.999f, 1e-3f);
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
