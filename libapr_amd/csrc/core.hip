/*
 * core.hip -- context, allocator, errors, geometry, B-spline coefficients and
 * the synthetic-volume generator of libaprh.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>

#include "aprh_internal.hpp"
#include "tiles.hpp"
#include "synth.hpp"

namespace {
thread_local std::string g_last_error;
}

namespace aprh {

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error %d (%s) at %s:%d in %s", (int)e, hipGetErrorString(e), file, line, what);
    int code = (e == hipErrorOutOfMemory) ? APRH_ERR_OUT_OF_MEMORY : APRH_ERR_HIP;
    throw Error(code, buf);
}

/* ---------------- Pool ---------------- */
size_t Pool::bucket(size_t b) {
    /* round up to 2 MiB granules above 2 MiB, else to a power of two (>= 256 B) */
    if (b <= 256) return 256;
    if (b <= (2u << 20)) {
        size_t p = 256;
        while (p < b) p <<= 1;
        return p;
    }
    const size_t g = 2u << 20;
    return (b + g - 1) / g * g;
}

void* Pool::get(size_t bytes) {
    size_t b = bucket(bytes);
    auto it = free_.lower_bound(b);
    /* reuse a cached block of the same bucket, or one at most 25% larger */
    if (it != free_.end() && it->first <= b + b / 4) {
        void* p = it->second;
        size_t sz = it->first;
        free_.erase(it);
        cached_ -= sz;
        live_[p] = sz;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        release_all();
        HIP_CHECK(hipMalloc(&p, b));
    }
    live_[p] = b;
    return p;
}

void Pool::put(void* p) {
    if (keep_) return;
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_.insert({it->second, p});
    cached_ += it->second;
    live_.erase(it);
}

void Pool::release_all() {
    for (auto& kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached_ = 0;
}

Pool::~Pool() {
    release_all();
    for (auto& kv : live_) (void)hipFree(kv.first);
}

/* ---------------- geometry ---------------- */
void geninfo_init(aprh_apr& a, int64_t y, int64_t x, int64_t z) {
    /* GenInfo::init (GenInfo.hpp:33-65) */
    a.org[0] = y; a.org[1] = x; a.org[2] = z;
    int64_t md = std::max(std::max(x, y), z);
    int lm = std::max(1, (int)std::ceil(std::log2((double)md)));
    a.l_min = 1;
    a.l_max = lm;
    a.y_num.assign(lm + 1, 0);
    a.x_num.assign(lm + 1, 0);
    a.z_num.assign(lm + 1, 0);
    for (int l = a.l_min; l <= a.l_max; ++l) {
        double cs = std::round(std::pow(2.0, (double)(a.l_max - l)));
        a.y_num[l] = (int64_t)std::ceil((double)y / cs);
        a.x_num[l] = (int64_t)std::ceil((double)x / cs);
        a.z_num[l] = (int64_t)std::ceil((double)z / cs);
    }
    /* LinearAccess::initialize_xz_linear (LinearAccess.hpp:37-54) */
    a.level_xz_vec.assign(a.l_max + 2, 0);
    uint64_t ct = 1;
    a.level_xz_vec[0] = 1;
    for (int i = 0; i <= a.l_max; ++i) {
        ct += (uint64_t)(a.x_num[i] * a.z_num[i]);
        a.level_xz_vec[i + 1] = ct;
    }
    a.n_xz = ct;
}

bool is_device_ptr(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

void copy_any(aprh_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
}

/* ---------------- B-spline coefficients ----------------
 * Smoothing-spline recursion constants (Unser 1993, eqs 4.5-4.8) and the
 * boundary initialisation vectors, evaluated with the reference's precision:
 * float operands, global double libm calls, float results
 * (ComputeGradient.hpp:256-330, impulse responses :213-223).            */
static float impulse_f(float k, float rho, float omg) {
    float ak = std::fabs(k);
    double num = std::pow((double)rho, (double)ak) * std::sin((double)((ak + 1.0f) * omg));
    return (float)(num / std::sin((double)omg));
}
static float impulse_b(float k, float rho, float omg, float gamma, float c0) {
    float ak = std::fabs(k);
    double p = std::pow((double)rho, (double)ak);
    double num = ((double)c0 * p) * (std::cos((double)(omg * ak)) + (double)gamma * std::sin((double)(omg * ak)));
    double den = std::pow(1.0 - 2.0 * (double)rho * std::cos((double)omg) + std::pow((double)rho, 2.0), 2.0);
    return (float)(num * (1.0 / den));
}

IIRCoeffs iir_coeffs(float lambda, float tol, int64_t len) {
    IIRCoeffs c;
    std::memset(&c, 0, sizeof(c));
    const float l24 = 24.0f * lambda;
    const float s144 = 3.0f + 144.0f * lambda;
    const float xi = (float)((double)(1.0f - 96.0f * lambda) + (double)l24 * std::sqrt((double)s144));
    const float ixi = 1.0f / xi;
    const double num = (double)(l24 - 1.0f) - std::sqrt((double)xi);
    const double inner = (double)ixi * ((double)(48.0f * lambda) + (double)l24 * std::sqrt((double)s144));
    const float rho = (float)((num / (double)l24) * std::sqrt(inner));
    const float omg = (float)std::atan(std::sqrt((double)(ixi * (144.0f * lambda - 1.0f))));
    const double rr = std::pow((double)rho, 2.0);
    const double cw = std::cos((double)omg);
    const float two_rho = 2.0f * rho;
    const float c0 = (float)((((1.0 + rr) / (1.0 - rr)) * (1.0 - (double)two_rho * cw + rr)) / (1.0 + (double)two_rho * cw + rr));
    const float gamma = (float)(((1.0 - rr) / (1.0 + rr)) * (1.0 / std::tan((double)omg)));
    c.b1 = (float)((double)two_rho * cw);
    c.b2 = (float)(-std::pow((double)rho, 2.0));
    c.nf = (float)std::pow(1.0 - 2.0 * (double)rho * cw + rr, 2.0);
    const int64_t k0 = (int64_t)std::ceil(std::fabs(std::log((double)tol) / std::log((double)rho)));
    if (k0 < 2 || k0 > 64) throw Error(APRH_ERR_UNSUPPORTED, "B-spline boundary length k0 outside [2,64] (lambda/tol)");
    const int64_t mL = std::min(k0, len);
    c.mL = (int)mL;
    std::vector<float> f(k0 + 3), b(k0 + 3);
    for (int64_t k = 0; k < k0 + 3; ++k) {
        f[k] = impulse_f((float)k, rho, omg);
        b[k] = impulse_b((float)k, rho, omg, gamma, c0);
    }
    std::vector<float> v1(k0, 0.0f), v2(k0, 0.0f), v3(k0, 0.0f), v4(k0, 0.0f);
    v1[1] = f[0];
    for (int64_t k = 0; k < k0; ++k) v1[k] += f[k + 1];
    for (int64_t k = 0; k < k0; ++k) v2[k] = f[k];
    v3[0] = b[1];
    for (int64_t k = 0; k + 1 < k0; ++k) v3[k + 1] += b[k] + b[k + 2];
    v4[0] = b[0];
    for (int64_t k = 1; k < k0; ++k) v4[k] += 2.0f * b[k];
    /* fold the taps beyond the line length into the last one */
    for (int64_t k = mL; k < k0; ++k) {
        v1[mL - 1] += v1[k];
        v2[mL - 1] += v2[k];
        v3[mL - 1] += v3[k];
        v4[mL - 1] += v4[k];
    }
    for (int64_t k = 0; k < mL; ++k) {
        c.bc1[k] = v1[k];
        c.bc2[k] = v2[k];
        c.bc3[k] = v3[k];
        c.bc4[k] = v4[k];
    }
    return c;
}

}  // namespace aprh

/* ---------------- zero fill ---------------- */
namespace aprh {
namespace {
struct ZeroWords {
    struct Args {
        uint32_t* p;
        size_t nw;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "zero"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < a.nw; i += (size_t)a.G * blockDim.x) a.p[i] = 0u;
    }
};
struct ZeroBytes {
    struct Args {
        uint8_t* p;
        size_t n;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "zero_bytes"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < a.n; i += (size_t)a.G * blockDim.x) a.p[i] = 0;
    }
};
/* device-to-device copy: 16-byte words when both ends and the size allow, else bytes */
struct CopyBytes {
    struct Args {
        void* dst;
        const void* src;
        size_t n;
        int wide;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "copy"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        const size_t stride = (size_t)a.G * blockDim.x;
        if (a.wide) {
            uint4* d = (uint4*)a.dst;
            const uint4* s = (const uint4*)a.src;
            for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < a.n / 16; i += stride) d[i] = s[i];
        } else {
            uint8_t* d = (uint8_t*)a.dst;
            const uint8_t* s = (const uint8_t*)a.src;
            for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < a.n; i += stride) d[i] = s[i];
        }
    }
};
}  // namespace

void launch_zero(aprh_context* ctx, void* p, size_t bytes) {
    if (!bytes) return;
    if (((uintptr_t)p & 3) == 0 && (bytes & 3) == 0) {
        const size_t nw = bytes / 4;
        const int64_t G = stride_tiles((int64_t)nw, 256);
        launch_tiles<ZeroWords>(ctx, {(uint32_t*)p, nw, G}, G);
    } else {
        const int64_t G = stride_tiles((int64_t)bytes, 256);
        launch_tiles<ZeroBytes>(ctx, {(uint8_t*)p, bytes, G}, G);
    }
}

void launch_copy(aprh_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    const int wide = (((uintptr_t)dst | (uintptr_t)src | bytes) & 15u) == 0;
    const int64_t G = stride_tiles((int64_t)(wide ? bytes / 16 : bytes), 256);
    launch_tiles<CopyBytes>(ctx, {dst, src, bytes, wide, G}, G);
}
}  // namespace aprh

/* ---------------- synthetic volumes ---------------- */
namespace aprh {

__global__ void k_synth_blobs(uint32_t* acc, int64_t y_num, int64_t x_num, int64_t z_num, uint64_t seed, int64_t n_blobs,
                              float smin, float smax, float amin, float amax) {
    const int64_t bi = blockIdx.x;
    if (bi >= n_blobs) return;
    synth::Blob b = synth::blob(seed, (uint64_t)bi, y_num, x_num, z_num, smin, smax, amin, amax);
    int64_t y0, y1, x0, x1, z0, z1;
    synth::box(b.cy, b.sigma, y_num, y0, y1);
    synth::box(b.cx, b.sigma, x_num, x0, x1);
    synth::box(b.cz, b.sigma, z_num, z0, z1);
    const int64_t ny = y1 - y0 + 1, nx = x1 - x0 + 1, nz = z1 - z0 + 1;
    if (ny <= 0 || nx <= 0 || nz <= 0) return;
    const int64_t total = ny * nx * nz;
    for (int64_t t = threadIdx.x; t < total; t += blockDim.x) {
        const int64_t yy = y0 + t % ny;
        const int64_t r = t / ny;
        const int64_t xx = x0 + r % nx;
        const int64_t zz = z0 + r / nx;
        uint32_t c = synth::contribution(b, yy, xx, zz);
        if (c) atomicAdd(&acc[(size_t)(zz * x_num * y_num + xx * y_num + yy)], c);
    }
}

__global__ void k_synth_finish(const uint32_t* acc, uint16_t* out, size_t n, uint64_t seed, float background, int noise) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float lam = background + (float)acc[i] * (1.0f / 64.0f);
        if (noise) {
            out[i] = synth::noisy(seed, (uint64_t)i, lam);
        } else {
            float k = lam + 0.5f;
            out[i] = (uint16_t)(uint32_t)(k > 65535.0f ? 65535.0f : k);
        }
    }
}

void synth_device(aprh_context* ctx, uint16_t* out, int64_t y, int64_t x, int64_t z, uint64_t seed, int64_t n_blobs, float smin,
                  float smax, float amin, float amax, float background, int noise) {
    const size_t n = (size_t)(y * x * z);
    Buf acc(ctx->pool, n * 4);
    HIP_CHECK(hipMemsetAsync(acc.ptr(), 0, n * 4, ctx->stream));
    if (n_blobs > 0) {
        hipLaunchKernelGGL(k_synth_blobs, dim3((unsigned)n_blobs), dim3(256), 0, ctx->stream, acc.as<uint32_t>(), y, x, z, seed,
                           n_blobs, smin, smax, amin, amax);
        HIP_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_synth_finish, dim3(4096), dim3(256), 0, ctx->stream, acc.as<uint32_t>(), out, n, seed, background, noise);
    HIP_LAUNCH_CHECK();
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace aprh

using namespace aprh;

namespace aprh {
void reset_times(aprh_context* ctx, int first, int last) {
    for (int i = first; i <= last; ++i) ctx->ev_marked[i] = false;
    for (int i = first; i < last; ++i) ctx->stage_ms[i] = 0.0f;
}
void collect_times(aprh_context* ctx, int first, int last) {
    if (!ctx->timing) return;
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int i = first; i < last; ++i) {
        int j = i + 1;
        while (j <= last && !ctx->ev_marked[j]) ++j;
        if (ctx->ev_marked[i] && j <= last) HIP_CHECK(hipEventElapsedTime(&ctx->stage_ms[i], ctx->ev[i], ctx->ev[j]));
    }
}
void set_last_error(const std::string& s) { g_last_error = s; }
}

extern "C" {

const char* aprh_last_error(void) { return g_last_error.c_str(); }
int aprh_version(void) { return 1; }

int aprh_context_create(int device, aprh_context** out) {
    APRH_TRY
    if (!out) throw Error(APRH_ERR_INVALID_ARGUMENT, "out is NULL");
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Error(APRH_ERR_INVALID_ARGUMENT, "no such device");
    HIP_CHECK(hipSetDevice(device));
    try {
        noalloc_init(device);   /* the no-alloc entry point's scratch address on this device */
    } catch (const Error&) {
        /* left unresolved: aprh_sample_image_u16_noalloc retries and reports the error */
    }
    aprh_context* c = new aprh_context();
    c->device = device;
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->pool = new Pool(device);
    for (int i = 0; i <= ST_COUNT; ++i) HIP_CHECK(hipEventCreate(&c->ev[i]));
    HIP_CHECK(hipMalloc(&c->d_flags, 256));
    *out = c;
    APRH_CATCH
}

int aprh_context_destroy(aprh_context* ctx) {
    APRH_TRY
    if (!ctx) return APRH_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i <= ST_COUNT; ++i) (void)hipEventDestroy(ctx->ev[i]);
    (void)hipFree(ctx->d_flags);
    delete ctx->pool;
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    APRH_CATCH
}

int aprh_context_set_keep_intermediates(aprh_context* ctx, int keep) {
    APRH_TRY
    if (!ctx) throw Error(APRH_ERR_INVALID_ARGUMENT, "ctx is NULL");
    ctx->keep_intermediates = keep;
    APRH_CATCH
}

int aprh_context_set_timing(aprh_context* ctx, int on) {
    APRH_TRY
    if (!ctx) throw Error(APRH_ERR_INVALID_ARGUMENT, "ctx is NULL");
    ctx->timing = on;
    APRH_CATCH
}

int aprh_context_set_conv_mode(aprh_context* ctx, int mode) {
    APRH_TRY
    if (!ctx) throw Error(APRH_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (mode != APRH_CONV_EXACT && mode != APRH_CONV_SEPARABLE) throw Error(APRH_ERR_INVALID_ARGUMENT, "unknown convolution mode");
    ctx->conv_mode = mode;
    APRH_CATCH
}

int aprh_context_stage_times(aprh_context* ctx, float* ms, int n) {
    APRH_TRY
    if (!ctx || !ms) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int i = 0; i < ST_COUNT && i < n; ++i) ms[i] = ctx->stage_ms[i];
    APRH_CATCH
}

const char* aprh_stage_name(int i) {
    static const char* names[ST_COUNT] = {"offset", "smooth_bspline_y", "smooth_bspline_x", "smooth_bspline_z",
                                          "gradient_ds", "inv_bspline", "local_intensity_scale", "levels",
                                          "pulling_scheme", "linear_structure", "sample_image"};
    return (i >= 0 && i < ST_COUNT) ? names[i] : "";
}

int aprh_stage_count(void) { return ST_COUNT; }

int aprh_context_release_cache(aprh_context* ctx) {
    APRH_TRY
    if (!ctx) throw Error(APRH_ERR_INVALID_ARGUMENT, "ctx is NULL");
    HIP_CHECK(hipSetDevice(ctx->device));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->pool->release_all();
    APRH_CATCH
}

int aprh_malloc(aprh_context* ctx, size_t bytes, void** out) {
    APRH_TRY
    if (!ctx || !out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    HIP_CHECK(hipSetDevice(ctx->device));
    HIP_CHECK(hipMalloc(out, bytes ? bytes : 1));
    APRH_CATCH
}

int aprh_free(aprh_context* ctx, void* p) {
    APRH_TRY
    if (!ctx) throw Error(APRH_ERR_INVALID_ARGUMENT, "ctx is NULL");
    HIP_CHECK(hipSetDevice(ctx->device));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    HIP_CHECK(hipFree(p));
    APRH_CATCH
}

int aprh_memcpy(aprh_context* ctx, void* dst, const void* src, size_t bytes) {
    APRH_TRY
    if (!ctx || (!dst && bytes) || (!src && bytes)) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    HIP_CHECK(hipSetDevice(ctx->device));
    copy_any(ctx, dst, src, bytes);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    APRH_CATCH
}

int aprh_synth_blobs_host(uint16_t* out, int64_t y, int64_t x, int64_t z, uint64_t seed, int64_t n_blobs, float smin, float smax,
                          float amin, float amax, float background, int noise) {
    APRH_TRY
    if (!out || y <= 0 || x <= 0 || z <= 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad synth arguments");
    const size_t n = (size_t)(y * x * z);
    std::vector<uint32_t> acc(n, 0);
    /* threads own z-slabs; the integer accumulation makes the bytes independent of the split */
    host_parallel_for(z, [&](int64_t zb, int64_t ze) {
        for (int64_t i = 0; i < n_blobs; ++i) {
            synth::Blob b = synth::blob(seed, (uint64_t)i, y, x, z, smin, smax, amin, amax);
            int64_t y0, y1, x0, x1, z0, z1;
            synth::box(b.cy, b.sigma, y, y0, y1);
            synth::box(b.cx, b.sigma, x, x0, x1);
            synth::box(b.cz, b.sigma, z, z0, z1);
            z0 = std::max(z0, zb);
            z1 = std::min(z1, ze - 1);
            for (int64_t zz = z0; zz <= z1; ++zz)
                for (int64_t xx = x0; xx <= x1; ++xx)
                    for (int64_t yy = y0; yy <= y1; ++yy) acc[(size_t)(zz * x * y + xx * y + yy)] += synth::contribution(b, yy, xx, zz);
        }
        for (size_t i = (size_t)(zb * x * y); i < (size_t)(ze * x * y); ++i) {
            float lam = background + (float)acc[i] * (1.0f / 64.0f);
            if (noise) {
                out[i] = synth::noisy(seed, (uint64_t)i, lam);
            } else {
                float k = lam + 0.5f;
                out[i] = (uint16_t)(uint32_t)(k > 65535.0f ? 65535.0f : k);
            }
        }
    });
    APRH_CATCH
}

int aprh_synth_blobs_device(aprh_context* ctx, uint16_t* out, int64_t y, int64_t x, int64_t z, uint64_t seed, int64_t n_blobs,
                            float smin, float smax, float amin, float amax, float background, int noise) {
    APRH_TRY
    if (!ctx || !out || y <= 0 || x <= 0 || z <= 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad synth arguments");
    HIP_CHECK(hipSetDevice(ctx->device));
    if (!is_device_ptr(out)) throw Error(APRH_ERR_INVALID_ARGUMENT, "out must be device memory");
    synth_device(ctx, out, y, x, z, seed, n_blobs, smin, smax, amin, amax, background, noise);
    APRH_CATCH
}

}  // extern "C"
