// gather_noalloc.hip — rzk_gather_dev_noalloc: the VPL gather of rzk_gather
// (visibility bits, then the shading sum) over device-resident receivers into a
// device-resident image, with no device allocation at all.  Its only scratch,
// the (VPL, 64-receiver group) visibility bit matrix, lives in the file-scope
// device array `scratch` below, so callers that must not allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc) can
// call it.
//
// `scratch` is a host-side stand-in in host code: kernels get the device
// address resolved by hipGetSymbolAddress for each device (gather_noalloc_setup, run
// by rzk_ctx_create with the device current), never the array's name.
#pragma clang fp contract(off)

#include <mutex>

#include "rzk_internal.h"

namespace {

// 32 MiB per device: ngroups x nvpl <= 4,194,304 words, e.g. 65,536 receivers x
// 4,096 VPLs or 262,144 receivers x 1,024 VPLs per call.
constexpr int64_t kScratchWords = (int64_t)4 << 20;
static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint64_t* g_scratch[kMaxDevices] = {};   // device address of `scratch`, per device

}  // namespace

// Resolves this device's copy of `scratch` (once per device; the device must be
// current).  Every GPU holds its own copy at its own address.
void gather_noalloc_setup(int device) {
    RZK_CHECK(device >= 0 && device < kMaxDevices, RZK_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    RZK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    RZK_CHECK(p != nullptr, RZK_E_HIP, "hipGetSymbolAddress returned null for the gather scratch");
    g_scratch[device] = static_cast<uint64_t*>(p);
}

static uint64_t* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

extern "C" int rzk_gather_dev_noalloc(rzk_ctx* ctx, const rzk_scene* s, const rzk_vpls* v, const float* d_recv,
                                      const int32_t* d_recv_mat, int64_t npts, int clamp_mode, float clamp,
                                      float tmin, float n_emitted, float* d_out_rgb) {
    if (!ctx) return RZK_E_INVALID;
    try {
        RZK_CHECK(s && v && npts >= 0 && (npts == 0 || (d_recv && d_recv_mat && d_out_rgb)), RZK_E_INVALID,
                  "bad gather arguments");
        RZK_CHECK(s->ctx->device == ctx->device && v->ctx->device == ctx->device, RZK_E_INVALID,
                  "scene, VPL set and context are on different devices");
        RZK_CHECK(clamp_mode >= 0 && clamp_mode <= 3, RZK_E_INVALID, "bad clamp mode");
        if (npts > 0) {
            RZK_CHECK(s->bvh.ntri > 0, RZK_E_INVALID, "empty scene");
            const int64_t ngroups = (npts + 63) / 64;
            const int64_t words = ngroups * v->n;
            RZK_CHECK(v->n == 0 || words / v->n == ngroups, RZK_E_INVALID, "gather size overflows");
            RZK_CHECK(words <= kScratchWords, RZK_E_INVALID,
                      "gather too large for the no-allocation scratch: ceil(npts / 64) x VPLs must be at most "
                      "4,194,304 (32 MiB of visibility bits)");
            uint64_t* bits = scratch_of(ctx->device);
            RZK_CHECK(bits != nullptr, RZK_E_STATE, "no-allocation scratch not set up on this device");
            // k_visibility writes every (VPL, group) word it covers, so the
            // scratch needs no clearing between calls
            SceneView sv = scene_view(s);
            launch_visibility(ctx, sv, d_recv, d_recv_mat, npts, nullptr, v->d_vpl, v->n, tmin, bits, nullptr);
            launch_gather_shade(ctx, sv, d_recv, d_recv_mat, npts, nullptr, v->d_vpl, v->d_mat, v->n, bits,
                                clamp_mode, clamp, n_emitted, nullptr, nullptr, d_out_rgb);
        }
        ctx->last_error.clear();
        return RZK_OK;
    } catch (const RzkError& e) {
        ctx->last_error = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        ctx->last_error = e.what();
        return RZK_E_INVALID;
    }
}
