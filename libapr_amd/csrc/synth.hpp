/*
 * synth.hpp -- deterministic synthetic volume generator (Gaussian blobs on a
 * background plus Poisson-like shot noise).
 *
 * Shared, bit-for-bit, by the host implementation and the device kernel:
 * only IEEE basic operations (+ - * / sqrt) and integer hashing are used, the
 * exponential is evaluated by the polynomial below, blob contributions are
 * accumulated in 32-bit fixed point (order independent), and the noise is a
 * pure function of (seed, voxel index).  Compiled with -ffp-contract=off the
 * host and the GPU therefore produce identical bytes.
 *
 * The blob model (centre uniform in the volume, sigma and amplitude uniform in
 * a range, truncated at 4 sigma) and the seeds are those of the benchmark
 * workloads.  The noise model draws Poisson(lambda) exactly by multiplication
 * of uniforms for lambda < 30 and by a rounded normal N(lambda, lambda) built
 * from 12 uniforms above that.
 */
#ifndef APRH_SYNTH_HPP
#define APRH_SYNTH_HPP

#include <stdint.h>

#ifdef __HIPCC__
#define SYN_HD __host__ __device__ inline
#else
#define SYN_HD inline
#endif

namespace synth {

SYN_HD uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform float in [0,1) from 24 random bits */
SYN_HD float uniform(uint64_t seed, uint64_t a, uint64_t b) {
    uint64_t h = mix64(seed ^ mix64(a * 0x100000001B3ull + b));
    return (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
}

/* 2^n for integer n in [-126, 127] */
SYN_HD float pow2i(int n) {
    uint32_t bits = (uint32_t)(n + 127) << 23;
    float f;
    __builtin_memcpy(&f, &bits, 4);
    return f;
}

/* exp(x) for x in [-87, 0]: x*log2(e) = n + f, 2^f by a degree-8 series of
   exp(f*ln2) on f in [0,1). */
SYN_HD float expneg(float x) {
    if (x < -87.0f) return 0.0f;
    float t = x * 1.44269504088896341f;
    int n = (int)t;
    if ((float)n > t) n -= 1;
    float f = (t - (float)n) * 0.693147180559945309f;
    float p = 1.0f + f * (1.0f + f * (0.5f + f * (0.166666667f + f * (0.0416666667f + f * (0.00833333333f +
              f * (0.00138888889f + f * (0.000198412698f + f * 0.0000248015873f)))))));
    return p * pow2i(n);
}

struct Blob {
    float cy, cx, cz, sigma, amp;
};

SYN_HD Blob blob(uint64_t seed, uint64_t i, int64_t y_num, int64_t x_num, int64_t z_num, float smin, float smax, float amin,
                 float amax) {
    Blob b;
    b.cy = uniform(seed, i, 0) * (float)y_num;
    b.cx = uniform(seed, i, 1) * (float)x_num;
    b.cz = uniform(seed, i, 2) * (float)z_num;
    b.sigma = smin + uniform(seed, i, 3) * (smax - smin);
    b.amp = amin + uniform(seed, i, 4) * (amax - amin);
    return b;
}

/* fixed-point contribution of blob b at voxel (y,x,z); 0 beyond 4 sigma */
SYN_HD uint32_t contribution(const Blob& b, int64_t y, int64_t x, int64_t z) {
    float dy = (float)y - b.cy, dx = (float)x - b.cx, dz = (float)z - b.cz;
    float d2 = dy * dy + dx * dx + dz * dz;
    float s2 = b.sigma * b.sigma;
    if (d2 > 16.0f * s2) return 0u;
    float e = expneg(-d2 / (2.0f * s2));
    return (uint32_t)(b.amp * e * 64.0f + 0.5f);
}

/* noisy sample of mean lambda for voxel index v */
SYN_HD uint16_t noisy(uint64_t seed, uint64_t v, float lam) {
    uint64_t s = mix64(seed + 0x5bd1e995ull);
    float k;
    if (lam < 30.0f) {
        float L = expneg(-lam);
        float p = 1.0f;
        int n = 0;
        for (int i = 0; i < 256; ++i) {
            p = p * uniform(s, v, (uint64_t)i);
            if (p <= L) break;
            n++;
        }
        k = (float)n;
    } else {
        float acc = 0.0f;
        for (int i = 0; i < 12; ++i) acc = acc + uniform(s, v, (uint64_t)i);
        float zz = acc - 6.0f;
        k = lam + __builtin_sqrtf(lam) * zz + 0.5f;
        k = k < 0.0f ? 0.0f : k;
    }
    if (k > 65535.0f) k = 65535.0f;
    return (uint16_t)(uint32_t)k;
}

/* integer bounding box of the 4-sigma ball along one axis */
SYN_HD void box(float c, float sigma, int64_t n, int64_t& lo, int64_t& hi) {
    float r = 4.0f * sigma;
    int64_t a = (int64_t)(c - r) - 1, b = (int64_t)(c + r) + 1;
    lo = a < 0 ? 0 : a;
    hi = b > n - 1 ? n - 1 : b;
}

}  // namespace synth

#endif
