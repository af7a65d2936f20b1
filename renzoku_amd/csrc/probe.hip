// probe.hip — libprobe.so: a vector-issue probe for the benchmark's rooflines,
// and the device evaluation of the samplers' fp32 transcendentals for the tests
// (not part of the engine).  k_valu_probe keeps sixteen independent FMA chains
// per lane, enough waves on every SIMD to hide the FMA latency, so it issues one
// VALU instruction per SIMD slot; its SQ_INSTS_VALU x 64 / launch time is the
// sustained VALU issue rate of the device at its running clock.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rzk_math.h"

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

// Separate multiply and add (the contraction-free dot products of the
// clustering), sixteen independent chains per lane.
__global__ __launch_bounds__(256) void k_valu_probe_muladd(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = gid * 1e-7f + (float)k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) a[k] = a[k] * m + c;
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += a[k];
    out[gid] = s;
}

// Minimum / maximum only, sixteen independent chains per lane.
__global__ __launch_bounds__(256) void k_valu_probe_minmax(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = gid * 1e-7f + (float)k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) a[k] = fmaxf(fminf(a[k], m), c);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += a[k];
    out[gid] = s;
}

// Multiply and add with every operand in a vector register (as in the
// clustering's dot products), thirty-two independent chains per lane.
__global__ __launch_bounds__(256) void k_valu_probe_vv(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float acc[8][4], mv[8], kv[4];
#pragma unroll
    for (int a = 0; a < 8; ++a) mv[a] = gid * 1e-9f + c * (float)a;
#pragma unroll
    for (int b = 0; b < 4; ++b) kv[b] = m - gid * 1e-9f * (float)b;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (float)(a + b);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = acc[a][b] * kv[b] + mv[a];
    }
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) s += acc[a][b];
    out[gid] = s;
}

// The slab-test mix of k_valu_probe_slab with every operand in a vector
// register.
__global__ __launch_bounds__(256) void k_valu_probe_slab_vv(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = gid * 1e-7f + 0.01f * (float)k;
    const float e = gid * 1e-12f;              // per-lane, so the constants below live in vector registers
    const float lo = -4.f + e, hi = 4.f - e, y = 2.5f + e, mv = m + e, cv = c - e;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            float t = (a[k] - cv) * mv;
            t = fmaxf(fminf(t, hi), lo);
            a[k] = t < y ? t : -t;
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += a[k];
    out[gid] = s;
}

// The slab mix with one scalar-register operand: the multiplier (a wave-uniform
// plane, as in the wave-coherent walks); the other operands in vector registers.
__global__ __launch_bounds__(256) void k_valu_probe_slab_sv(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = gid * 1e-7f + 0.01f * (float)k;
    const float e = gid * 1e-12f;
    const float lo = -4.f + e, hi = 4.f - e, y = 2.5f + e, cv = c - e;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            float t = (a[k] - cv) * m;
            t = fmaxf(fminf(t, hi), lo);
            a[k] = t < y ? t : -t;
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += a[k];
    out[gid] = s;
}

// The slab mix with scalar operands in the subtract and the multiply only.
__global__ __launch_bounds__(256) void k_valu_probe_slab_ss(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = gid * 1e-7f + 0.01f * (float)k;
    const float e = gid * 1e-12f;
    const float lo = -4.f + e, hi = 4.f - e, y = 2.5f + e;
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

// The slab mix with constants (inline and literal) in min, max and compare,
// the subtract and multiply on vector registers.
__global__ __launch_bounds__(256) void k_valu_probe_slab_kc(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = gid * 1e-7f + 0.01f * (float)k;
    const float e = gid * 1e-12f;
    const float mv = m + e, cv = c - e;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            float t = (a[k] - cv) * mv;
            t = fmaxf(fminf(t, 4.f), -4.f);
            a[k] = t < 2.5f ? t : -t;
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += a[k];
    out[gid] = s;
}

// FMA chains with every operand in a vector register.
__global__ __launch_bounds__(256) void k_valu_probe_fma_vv(float* __restrict__ out, int iters, float m, float c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = gid * 1e-7f + (float)k;
    const float e = gid * 1e-12f;
    const float mv = m + e, cv = c - e;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) a[k] = __builtin_fmaf(a[k], mv, cv);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += a[k];
    out[gid] = s;
}

// One launch of probe `kind` (0 FMA chains, 1 slab mix, 2 multiply + add,
// 3 min / max, 4 multiply + add on vector registers only, 5 the slab mix and
// 6 FMA chains with vector-register operands, 7 the slab mix with a scalar
// multiplier, 8 with a scalar subtrahend and multiplier, 9 with constant
// min / max / compare operands) over `blocks` x 256
// threads running `iters` rounds; the
// launch is timed with HIP events (after one untimed warm launch).  Returns 0,
// or the HIP error code.
extern "C" int rzk_probe_valu(int kind, int blocks, int iters, float* out_ms) {
    if (blocks <= 0 || iters <= 0 || !out_ms || kind < 0 || kind > 9) return (int)hipErrorInvalidValue;
    auto kern = kind == 0   ? k_valu_probe
                : kind == 1 ? k_valu_probe_slab
                : kind == 2 ? k_valu_probe_muladd
                : kind == 3 ? k_valu_probe_minmax
                : kind == 4 ? k_valu_probe_vv
                : kind == 5 ? k_valu_probe_slab_vv
                : kind == 6 ? k_valu_probe_fma_vv
                : kind == 7 ? k_valu_probe_slab_sv
                : kind == 8 ? k_valu_probe_slab_ss
                            : k_valu_probe_slab_kc;
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

// fn 0 sin, 1 cos, 2 tan, 3 atan, 4 atan2(x, y), 5 acos, 6 log, 7 exp, 8 pow(x, y) of rzk_math.h
__global__ void k_xmath(int fn, int n, const float* __restrict__ x, const float* __restrict__ y,
                        float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = x[i], b = y ? y[i] : 0.0f;
    float v;
    switch (fn) {
    case 0: v = rzk::xsin(a); break;
    case 1: v = rzk::xcos(a); break;
    case 2: v = rzk::xtan(a); break;
    case 3: v = rzk::xatan(a); break;
    case 4: v = rzk::xatan2(a, b); break;
    case 5: v = rzk::xacos(a); break;
    case 6: v = rzk::xlog(a); break;
    case 7: v = rzk::xexp(a); break;
    default: v = rzk::xpow(a, b); break;
    }
    out[i] = v;
}

// host arrays in and out (y may be NULL)
extern "C" int rzk_probe_xmath(int fn, int n, const float* x, const float* y, float* out) {
    if (n <= 0 || !x || !out || fn < 0 || fn > 8) return (int)hipErrorInvalidValue;
    float *dx = nullptr, *dy = nullptr, *dout = nullptr;
    const size_t bytes = sizeof(float) * (size_t)n;
    hipError_t err = hipMalloc((void**)&dx, bytes);
    if (err == hipSuccess) err = hipMalloc((void**)&dout, bytes);
    if (err == hipSuccess && y) err = hipMalloc((void**)&dy, bytes);
    if (err == hipSuccess) err = hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice);
    if (err == hipSuccess && y) err = hipMemcpy(dy, y, bytes, hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_xmath, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, fn, n, dx, dy, dout);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (dout) (void)hipFree(dout);
    return (int)err;
}
