/*
 * capi_noalloc.hip -- ParticleData<uint16_t>::sample_image without any device allocation.
 *
 * aprh_sample_image_u16 takes its mean pyramid (PixelData::downsamplePyramid,
 * PixelData.hpp:1036-1058) from the context's caching pool, which allocates on a
 * first call of a new size.  Callers that cannot allocate (code under stream capture,
 * serving loops, harnesses that forbid hipMalloc) use aprh_sample_image_u16_noalloc:
 * the same kernels, with the pyramid in `scratch`, a file-scope device array of
 * SCRATCH_BYTES.  Every call on one device shares that device's copy of the array,
 * so calls that overlap on one device race.  The pyramid levels are written in full
 * before the gather reads them, so nothing has to start at zero.
 */
#include <mutex>

#include "aprh_internal.hpp"

namespace {

constexpr size_t SCRATCH_BYTES = size_t(64) << 20;   /* the pyramid of a u16 image of up to ~235M voxels */
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) uint16_t scratch[SCRATCH_BYTES / sizeof(uint16_t)];

std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};   /* device address of `scratch` on each device (never the host symbol) */

}  // namespace

namespace aprh {

/* resolves `scratch` on the current device (called by aprh_context_create with `device` current) */
void noalloc_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw Error(APRH_ERR_UNSUPPORTED, "device index beyond the no-alloc scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(APRH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
}

/* the device address of `scratch` on `device`, which must be current */
static void* scratch_ptr(int device) {
    noalloc_init(device);   /* no-op once resolved */
    std::lock_guard<std::mutex> lk(g_mu);
    return g_scratch[device];
}

}  // namespace aprh

using namespace aprh;

extern "C" {

size_t aprh_noalloc_scratch_bytes(void) { return SCRATCH_BYTES; }

int aprh_sample_image_u16_noalloc(aprh_context* ctx, const aprh_apr* a, const uint16_t* image, uint16_t* parts) {
    APRH_TRY
    check_ctx(ctx);
    if (!a || !image || !parts) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!is_device_ptr(image) || !is_device_ptr(parts))
        throw Error(APRH_ERR_INVALID_ARGUMENT, "image and parts must be device memory (nothing is staged here)");
    const size_t need = sample_pyramid_bytes(*a, sizeof(uint16_t));
    if (need > SCRATCH_BYTES)
        throw Error(APRH_ERR_UNSUPPORTED, "mean pyramid of " + std::to_string(need) + " bytes exceeds the " +
                                               std::to_string(SCRATCH_BYTES) + "-byte no-alloc scratch");
    launch_sample_u16(ctx, *a, image, parts, (uint16_t*)scratch_ptr(ctx->device));
    APRH_CATCH
}

}  // extern "C"
