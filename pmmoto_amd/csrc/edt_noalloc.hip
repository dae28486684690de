// edt_noalloc.hip -- the exact EDT for callers that cannot allocate.
//
// pmx_edt_noalloc runs pmx_edt's kernels with their workspace in a
// file-scope device array instead of caller or library allocations, for code
// under HIP graph capture, tight serving loops and harnesses that forbid
// hipMalloc.  Every GPU the library runs on holds its own copy of the array
// (PMX_EDT_NOALLOC_SCRATCH_BYTES) until the process ends.
//
// The array's name is a host-side stand-in for the device symbol: its device
// address comes from hipGetSymbolAddress, once per device on that device's
// first call, and only that address is handed to the kernels.
//
// Nothing in the scratch needs to start at zero: pass Z resets the counters
// and the deferred-line bitmaps at the start of every EDT.  Every call on a
// GPU shares that GPU's array, so calls that overlap on one GPU (other
// streams, other host threads) race and must be serialised by the caller.
#include <mutex>

#include "edt_internal.hpp"

namespace pmx {
namespace {

constexpr size_t kScratchBytes = PMX_EDT_NOALLOC_SCRATCH_BYTES;
// fewer persistent fallback lanes than pmx_edt (their stacks live in the
// scratch); deferred lines are rare in porous media and the lanes loop
constexpr int kNoallocFallbackLanes = 2048;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// the current device's copy of `scratch`
int scratch_address(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(PMX_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(PMX_ERR_INVALID, "device %d beyond %d devices", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return fail(PMX_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        if (!p) return fail(PMX_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PMX_OK;
}

}  // namespace
}  // namespace pmx

extern "C" {

int pmx_edt_noalloc(const uint8_t *img, const pmx_geom *g, int32_t *d2, float *f32sq, float *f32, double *f64,
                    hipStream_t stream)
{
    using namespace pmx;
    clear_error();
    if (!d2 && !f32sq && !f32 && !f64) return fail(PMX_ERR_INVALID, "no output requested");
    if (int e = validate_geom(g)) return e;
    EdtPlan p;
    const size_t need = plan_edt(g, nullptr, &p, kNoallocFallbackLanes);
    if (need > kScratchBytes)
        return fail(PMX_ERR_INVALID,
                    "box %lld x %lld x %lld needs %zu bytes of scratch; pmx_edt_noalloc holds %zu (use pmx_edt)",
                    (long long)g->n[0], (long long)g->n[1], (long long)g->n[2], need, kScratchBytes);
    void *ws = nullptr;
    if (int e = scratch_address(&ws)) return e;
    PassXOut o{};
    o.d2 = d2;
    o.f32sq = f32sq;
    o.f32 = f32;
    o.f64 = f64;
    o.sx = g->out_s[0];
    o.sy = g->out_s[1];
    return run_edt(img, g, false, o, INT64_MAX, ws, kScratchBytes, stream, nullptr, 0, -1, kNoallocFallbackLanes);
}

}  // extern "C"
