/*
 * Host checks of the exact-arithmetic shortcuts used in the HIP kernels, restated in C with the
 * same operations (compiled with -ffp-contract=off; fmaf is a single rounding on x86-64 as on
 * gfx950):
 *   div_odd<D>(x)  = copysign(fma(fma(-q0, D, x), r, q0), x), q0 = x * r, r = RN(1/D)
 *                    must equal x / (float)D for every finite x, D = 1, 3, ..., 13
 *                    (k_grad_lis.hip, running means of the local intensity scale);
 *   isqrt_k(k)      = floor(sqrt(k)) from a float square root estimate and one integer
 *                    correction, for sums of three squares of differences of u16 values
 *                    (k_grad_lis.hip, gradient magnitude), against the reference's
 *                    trunc(sqrt(double)) form; the estimate is taken one ulp low, exact and
 *                    one ulp high (the hardware square root is within one ulp), and the
 *                    maximum over a cell is taken before the root (both maps are monotone).
 * Usage: arith_check [stride]   -- tests every stride-th float bit pattern (1 = all 2^32).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float div_odd(float x, float d, float r) {
    const float q0 = x * r;
    const float e = fmaf(-q0, d, x);
    return copysignf(fmaf(e, r, q0), x);
}

static uint32_t isqrt_k(uint64_t k, int ulp) {
    float r = sqrtf((float)k);
    if (ulp < 0) r = nextafterf(r, 0.0f);
    if (ulp > 0) r = nextafterf(r, INFINITY);
    uint32_t s = (uint32_t)r;
    if ((uint64_t)s * s > k) --s;
    else if ((uint64_t)(s + 1) * (s + 1) <= k) ++s;
    return s;
}
static uint32_t isqrt_d2(uint32_t a, uint32_t b, uint32_t c, int ulp) { return isqrt_k((uint64_t)a + b + c, ulp); }

static uint64_t rng = 88172645463325252ull;
static uint32_t next32(void) {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (uint32_t)(rng >> 16);
}

int main(int argc, char** argv) {
    const uint64_t stride = argc > 1 ? strtoull(argv[1], NULL, 10) : 97;
    uint64_t bad = 0, n = 0;
    for (int D = 1; D <= 13; D += 2) {
        const float d = (float)D, r = 1.0f / d;
        for (uint64_t u = 0; u < (1ull << 32); u += stride) {
            const uint32_t bits = (uint32_t)u;
            float x;
            memcpy(&x, &bits, 4);
            if (!isfinite(x)) continue;
            const float q = x / d, g = div_odd(x, d, r);
            ++n;
            if (memcmp(&q, &g, 4) != 0) {
                if (bad < 5) printf("div_odd D=%d x=%a: %a != %a\n", D, x, g, q);
                ++bad;
            }
        }
    }
    /* signed zeros, the smallest and largest magnitudes */
    const uint32_t special[8] = {0x00000000u, 0x80000000u, 0x00000001u, 0x80000001u, 0x00800000u, 0x80800000u, 0x7F7FFFFFu, 0xFF7FFFFFu};
    for (int D = 1; D <= 13; D += 2)
        for (int i = 0; i < 8; ++i) {
            float x;
            memcpy(&x, &special[i], 4);
            const float q = x / (float)D, g = div_odd(x, (float)D, 1.0f / (float)D);
            ++n;
            if (memcmp(&q, &g, 4) != 0) {
                if (bad < 5) printf("div_odd D=%d x=%a: %a != %a\n", D, x, g, q);
                ++bad;
            }
        }
    printf("div_odd: %llu values, %llu mismatches\n", (unsigned long long)n, (unsigned long long)bad);
    uint64_t bad2 = 0, n2 = 0;
    for (uint64_t i = 0; i < 20000000ull; ++i) {
        int dx = (int)(next32() % 131071u) - 65535, dy = (int)(next32() % 131071u) - 65535, dz = (int)(next32() % 131071u) - 65535;
        if (i & 1) {
            dx >>= next32() % 17;
            dy >>= next32() % 17;
            dz >>= next32() % 17;
        }
        if (i % 7 == 0) {   /* near perfect squares */
            dy = (int)(next32() % 4) - 2;
            dz = (int)(next32() % 4);
        }
        const uint32_t ux = (uint32_t)dx, uy = (uint32_t)dy, uz = (uint32_t)dz;
        const double q = (double)dx * dx / 4.0 + (double)dz * dz / 4.0 + (double)dy * dy / 4.0;
        const uint32_t ref = (uint32_t)sqrt(q);
        for (int ulp = -1; ulp <= 1; ++ulp) {
            const uint32_t g = isqrt_d2(ux * ux, uy * uy, uz * uz, ulp) >> 1;
            ++n2;
            if (g != ref) {
                if (bad2 < 5) printf("isqrt_k %d %d %d (ulp %d): %u != %u\n", dx, dy, dz, ulp, g, ref);
                ++bad2;
            }
        }
    }
    /* the extremes: every square sum of three u16 differences from 0 to 65535 */
    const int ext[4] = {0, 1, 65534, 65535};
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int c = 0; c < 4; ++c) {
                const uint32_t ux = (uint32_t)ext[a], uy = (uint32_t)ext[b], uz = (uint32_t)ext[c];
                const uint32_t g = isqrt_d2(ux * ux, uy * uy, uz * uz, 0) >> 1;
                const double q = (double)ext[a] * ext[a] / 4.0 + (double)ext[c] * ext[c] / 4.0 + (double)ext[b] * ext[b] / 4.0;
                ++n2;
                if (g != (uint32_t)sqrt(q)) ++bad2;
            }
    /* a cell's maximum: the root of the largest sum equals the largest root (8 random voxels) */
    for (uint64_t i = 0; i < 2000000ull; ++i) {
        uint64_t kmax = 0;
        uint32_t gmax = 0;
        for (int v = 0; v < 8; ++v) {
            const int dx = (int)(next32() % 131071u) - 65535, dy = (int)(next32() % 2001u) - 1000, dz = (int)(next32() % 257u) - 128;
            const uint32_t ux = (uint32_t)dx, uy = (uint32_t)dy, uz = (uint32_t)dz;
            const uint64_t k = (uint64_t)(ux * ux) + (uint64_t)(uy * uy) + (uint64_t)(uz * uz);
            kmax = k > kmax ? k : kmax;
            const double q = (double)dx * dx / 4.0 + (double)dz * dz / 4.0 + (double)dy * dy / 4.0;
            const uint32_t g = (uint32_t)sqrt(q);
            gmax = g > gmax ? g : gmax;
        }
        ++n2;
        if ((isqrt_k(kmax, 0) >> 1) != gmax) ++bad2;
    }
    printf("isqrt_k: %llu values, %llu mismatches\n", (unsigned long long)n2, (unsigned long long)bad2);
    /* range check of the y smoothing's first sweep: roundf(t) lies in [0, 65535] exactly when
       -0.5 < t < 65535.5 (every float near both bounds, and every stride-th bit pattern) */
    uint64_t n3 = 0, bad3 = 0;
    const float lim[2][2] = {{-0.75f, -0.25f}, {65533.0f, 65538.0f}};
    for (int r = 0; r < 2; ++r)
        for (float t = lim[r][0]; t <= lim[r][1]; t = nextafterf(t, INFINITY)) {
            const float q = roundf(t);
            ++n3;
            if ((q >= 0.0f && q <= 65535.0f) != (t > -0.5f && t < 65535.5f)) ++bad3;
        }
    for (uint64_t b = 0; b < (1ull << 32); b += stride) {
        const uint32_t u = (uint32_t)b;
        float t;
        memcpy(&t, &u, 4);
        if (t != t) continue;
        const float q = roundf(t);
        ++n3;
        if ((q >= 0.0f && q <= 65535.0f) != (t > -0.5f && t < 65535.5f)) ++bad3;
    }
    printf("u16 range: %llu values, %llu mismatches\n", (unsigned long long)n3, (unsigned long long)bad3);
    return (bad || bad2 || bad3) ? 1 : 0;
}
