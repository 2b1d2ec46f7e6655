// hm_scan_noalloc: the scan.wgsl dispatch over caller-owned device buffers,
// allocating nothing. Its one scratch buffer, the per-bundle exclusive
// prefix that k_scan writes and k_expand searches, is the file-scope device
// array `scratch` below. Every GPU the library runs on holds its own copy
// until the process ends; hm_create looks up that copy's address once per
// device (hipGetSymbolAddress, with the device current) and the entry point
// passes that address, never the array's name, to the kernels.
#include <mutex>

#include "hm_ctx.h"
#include "hm_kernels.h"

static __device__ __align__(256) uint32_t scratch[HM_SCAN_NOALLOC_MAX_BUNDLES + 1u];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device

void* scratch_for(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mutex);
    return g_scratch[device];
}

}  // namespace

// Looks up the address of `scratch` on `device`, which must be the current
// device; once per device (hm_create calls it).
hipError_t hm_noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[device] = p;
    return hipSuccess;
}

extern "C" hm_status hm_scan_noalloc(hm_ctx* ctx, const hm_lod_record* lod, uint32_t bundle_count,
                                     hm_strand_ref* refs, uint32_t capacity, uint32_t* indirect) {
    if (!ctx) return HM_ERR_INVALID;
    if (!indirect || (bundle_count && !lod) || (capacity && !refs))
        return hm_fail(ctx, HM_ERR_INVALID, "hm_scan_noalloc: null buffer");
    if (bundle_count > HM_SCAN_NOALLOC_MAX_BUNDLES)
        return hm_fail(ctx, HM_ERR_CAPACITY,
                       "hm_scan_noalloc: " + std::to_string(bundle_count) + " bundles exceed the maximum of " +
                           std::to_string(HM_SCAN_NOALLOC_MAX_BUNDLES));
    void* prefix = scratch_for(ctx->device);
    if (!prefix) return hm_fail(ctx, HM_ERR_NOT_READY, "hm_scan_noalloc: scratch address not set up for the device");
    // k_scan writes prefix[0 .. bundle_count] whole on every call, so the
    // array needs no clearing between calls.
    const hipError_t e = hm::launch_scan(ctx->stream, lod, bundle_count, capacity, static_cast<uint32_t*>(prefix),
                                         indirect, refs);
    if (e != hipSuccess) return hm_hip_fail(ctx, e, "hm_scan_noalloc: launch_scan");
    return HM_OK;
}
