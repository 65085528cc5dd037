/*
 * k_synth.hip -- synthetic deconvolution workloads (benchmark / test inputs).
 *
 * A ground-truth volume of Gaussian blobs on a background (the blob model of
 * synth.hpp, accumulated in fixed point so the bytes do not depend on the
 * order of the atomics), blurred by a separable Gaussian PSF with reflective
 * boundaries, then shot noise drawn per voxel (synth::noisy) into uint16.
 */
#include <cmath>

#include "aprh_internal.hpp"
#include "synth.hpp"

namespace aprh {

namespace {

__global__ void k_blobs_acc(uint32_t* acc, int64_t y_num, int64_t x_num, int64_t z_num, uint64_t seed, int64_t n_blobs, float smin,
                            float smax, float amin, float amax) {
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

__global__ void k_acc_to_f32(const uint32_t* __restrict__ acc, float* out, size_t n, float background) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = background + (float)acc[i] * (1.0f / 64.0f);
}

__device__ __forceinline__ int64_t reflect(int64_t q, int64_t n) {
    if (q < 0) q = -q;
    if (q >= n) q = 2 * (n - 1) - q;
    return q < 0 ? 0 : (q >= n ? n - 1 : q);
}

/* one separable pass: out = sum_k w[k] * in(p + (k - h) e_axis), reflective ends */
__global__ void k_blur_axis(const float* __restrict__ in, float* out, int64_t yn, int64_t xn, int64_t zn, int axis,
                            const float* __restrict__ w, int taps) {
    const size_t n = (size_t)(yn * xn * zn);
    const int h = taps / 2;
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? yn : yn * xn);
    const int64_t len = axis == 0 ? yn : (axis == 1 ? xn : zn);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int64_t c = axis == 0 ? (int64_t)(i % (size_t)yn)
                                    : (axis == 1 ? (int64_t)((i / (size_t)yn) % (size_t)xn) : (int64_t)(i / (size_t)(yn * xn)));
        const float* base = in + (int64_t)i - c * stride;
        float acc = 0.0f;
        for (int k = 0; k < taps; ++k) acc += w[k] * base[reflect(c + k - h, len) * stride];
        out[i] = acc;
    }
}

__global__ void k_noise_u16(const float* __restrict__ lam, uint16_t* out, size_t n, uint64_t seed) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float l = lam[i] < 0.0f ? 0.0f : lam[i];
        out[i] = synth::noisy(seed, (uint64_t)i, l);
    }
}

/* normalised 1-D Gaussian factors of the PSF, (y, x, z) one after the other */
std::vector<float> psf_factors(const float* psf_sigma3, int taps) {
    std::vector<float> w(3 * (size_t)taps);
    for (int d = 0; d < 3; ++d) {
        double s = 0;
        for (int k = 0; k < taps; ++k) {
            const double t = k - taps / 2;
            const double v = psf_sigma3[d] > 0 ? std::exp(-t * t / (2.0 * psf_sigma3[d] * psf_sigma3[d])) : (t == 0 ? 1.0 : 0.0);
            w[(size_t)d * taps + k] = (float)v;
            s += v;
        }
        for (int k = 0; k < taps; ++k) w[(size_t)d * taps + k] = (float)(w[(size_t)d * taps + k] / s);
    }
    return w;
}

int64_t reflect_host(int64_t q, int64_t n) {
    if (q < 0) q = -q;
    if (q >= n) q = 2 * (n - 1) - q;
    return q < 0 ? 0 : (q >= n ? n - 1 : q);
}

}  // namespace

/* the same workload on the host, operation for operation (blob sums in fixed point, each
   blur tap a float multiply then add in tap order, per-voxel noise), so the bytes equal the
   device's; threads own z-slabs */
void synth_deconv_host(float* gt, uint16_t* noisy, int64_t y, int64_t x, int64_t z, uint64_t seed, int64_t n_blobs, float smin,
                       float smax, float amin, float amax, float background, const float* psf_sigma3, int taps) {
    const size_t n = (size_t)(y * x * z);
    const size_t plane = (size_t)(y * x);
    {
        std::vector<uint32_t> acc(n, 0);
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
            for (size_t i = (size_t)zb * plane; i < (size_t)ze * plane; ++i) gt[i] = background + (float)acc[i] * (1.0f / 64.0f);
        });
    }
    const std::vector<float> w = psf_factors(psf_sigma3, taps);
    const int h = taps / 2;
    std::vector<float> a(n), b(n);
    auto pass = [&](const float* in, float* out, int axis) {
        const int64_t stride = axis == 0 ? 1 : (axis == 1 ? y : y * x);
        const int64_t len = axis == 0 ? y : (axis == 1 ? x : z);
        const float* wk = w.data() + (size_t)axis * taps;
        host_parallel_for(z, [&](int64_t zb, int64_t ze) {
            for (int64_t zz = zb; zz < ze; ++zz)
                for (int64_t xx = 0; xx < x; ++xx)
                    for (int64_t yy = 0; yy < y; ++yy) {
                        const int64_t i = (zz * x + xx) * y + yy;
                        const int64_t c = axis == 0 ? yy : (axis == 1 ? xx : zz);
                        const float* base = in + i - c * stride;
                        float s = 0.0f;
                        for (int k = 0; k < taps; ++k) s += wk[k] * base[reflect_host(c + k - h, len) * stride];
                        out[i] = s;
                    }
        });
    };
    pass(gt, a.data(), 0);
    pass(a.data(), b.data(), 1);
    pass(b.data(), a.data(), 2);
    host_parallel_for(z, [&](int64_t zb, int64_t ze) {
        for (size_t i = (size_t)zb * plane; i < (size_t)ze * plane; ++i) {
            const float l = a[i] < 0.0f ? 0.0f : a[i];
            noisy[i] = synth::noisy(seed, (uint64_t)i, l);
        }
    });
}

void synth_deconv_device(aprh_context* ctx, float* gt, uint16_t* noisy, int64_t y, int64_t x, int64_t z, uint64_t seed, int64_t n_blobs,
                         float smin, float smax, float amin, float amax, float background, const float* psf_sigma3, int taps) {
    hipStream_t st = ctx->stream;
    const size_t n = (size_t)(y * x * z);
    {
        Buf acc(ctx->pool, n * 4);
        HIP_CHECK(hipMemsetAsync(acc.ptr(), 0, n * 4, st));
        if (n_blobs > 0) {
            hipLaunchKernelGGL(k_blobs_acc, dim3((unsigned)n_blobs), dim3(256), 0, st, acc.as<uint32_t>(), y, x, z, seed, n_blobs, smin,
                               smax, amin, amax);
            HIP_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_acc_to_f32, dim3(4096), dim3(256), 0, st, acc.as<uint32_t>(), gt, n, background);
        HIP_LAUNCH_CHECK();
    }
    const std::vector<float> w = psf_factors(psf_sigma3, taps);
    Buf dw(ctx->pool, w.size() * 4), a(ctx->pool, n * 4), b(ctx->pool, n * 4);
    HIP_CHECK(hipMemcpyAsync(dw.ptr(), w.data(), w.size() * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_blur_axis, dim3(8192), dim3(256), 0, st, (const float*)gt, a.as<float>(), y, x, z, 0, dw.as<float>(), taps);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_blur_axis, dim3(8192), dim3(256), 0, st, (const float*)a.as<float>(), b.as<float>(), y, x, z, 1,
                       dw.as<float>() + taps, taps);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_blur_axis, dim3(8192), dim3(256), 0, st, (const float*)b.as<float>(), a.as<float>(), y, x, z, 2,
                       dw.as<float>() + 2 * taps, taps);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_noise_u16, dim3(8192), dim3(256), 0, st, (const float*)a.as<float>(), noisy, n, seed);
    HIP_LAUNCH_CHECK();
    HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace aprh

extern "C" int aprh_synth_deconv_device(aprh_context* ctx, float* gt, uint16_t* noisy, int64_t y, int64_t x, int64_t z, uint64_t seed,
                                        int64_t n_blobs, float sigma_min, float sigma_max, float amp_min, float amp_max,
                                        float background, const float* psf_sigma3, int32_t psf_taps) {
    APRH_TRY
    aprh::check_ctx(ctx);
    if (!gt || !noisy || !psf_sigma3 || y <= 0 || x <= 0 || z <= 0 || psf_taps < 1 || psf_taps % 2 == 0 || psf_taps > 63)
        throw aprh::Error(APRH_ERR_INVALID_ARGUMENT, "bad synth arguments");
    if (!aprh::is_device_ptr(gt) || !aprh::is_device_ptr(noisy)) throw aprh::Error(APRH_ERR_INVALID_ARGUMENT, "outputs must be device memory");
    aprh::synth_deconv_device(ctx, gt, noisy, y, x, z, seed, n_blobs, sigma_min, sigma_max, amp_min, amp_max, background, psf_sigma3,
                              psf_taps);
    APRH_CATCH
}

extern "C" int aprh_synth_deconv_host(float* gt, uint16_t* noisy, int64_t y, int64_t x, int64_t z, uint64_t seed, int64_t n_blobs,
                                      float sigma_min, float sigma_max, float amp_min, float amp_max, float background,
                                      const float* psf_sigma3, int32_t psf_taps) {
    APRH_TRY
    if (!gt || !noisy || !psf_sigma3 || y <= 0 || x <= 0 || z <= 0 || psf_taps < 1 || psf_taps % 2 == 0 || psf_taps > 63)
        throw aprh::Error(APRH_ERR_INVALID_ARGUMENT, "bad synth arguments");
    aprh::synth_deconv_host(gt, noisy, y, x, z, seed, n_blobs, sigma_min, sigma_max, amp_min, amp_max, background, psf_sigma3,
                            psf_taps);
    APRH_CATCH
}
