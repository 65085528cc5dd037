/*
 * k_iir.hip -- B-spline smoothing of the conversion pipeline on gfx950.
 *
 *  - min/max reduction + bspline offset     (AutoParameters.hpp:165-196)
 *  - y pass (ComputeGradient.hpp:248-413): one lane per y-line.  u16 lines of
 *    a multiple of 64 samples go through an LDS tile (64 lines x 64 samples)
 *    so global accesses stay coalesced, and do not store the causal result
 *    (per-chunk recursion states are saved and the chunk is recomputed for the
 *    anti-causal sweep); other lines whose length is a multiple of the 16-byte
 *    vector are streamed with 16-byte loads/stores, several vectors ahead of
 *    the recursion.  The offset copy of computeL (APRConverter.hpp:213-224) is
 *    fused into the first read.
 *  - x / z passes (ComputeGradient.hpp:415-599, 601-779): lanes run along y
 *    (8 u16 / 4 f32 samples per lane), so each recursion step is one
 *    coalesced 16-byte-per-lane row access.  Axes of 64 samples or more save
 *    the causal state every 16 steps instead of storing the causal result and
 *    recompute each chunk in the anti-causal sweep, one chunk loaded ahead.
 *
 * Every lane keeps the reference's per-line FP32 operation order; the build
 * disables FP contraction, so results are bit-identical to the CPU path.
 */
#include <type_traits>

#include "aprh_internal.hpp"
#include "tiles.hpp"

namespace aprh {

namespace {

__device__ __forceinline__ uint16_t f2u16(float v) {
    /* g++/x86-64 float->uint16 conversion: truncation through int32 */
    if (!(v > -2147483648.0f && v < 2147483648.0f)) return 0;
    return (uint16_t)(int32_t)v;
}

/* roundf (half away from zero) as trunc(v + copysign(0.5 - 2^-25, v)): three VALU operations,
   and equal to roundf bit for bit on every non-NaN float (v + 0.49999997 crosses the next
   integer exactly when the fraction of |v| is at least one half; checked over all 2^32
   patterns against the C library) */
__device__ __forceinline__ float round_away(float v) { return truncf(v + copysignf(0.49999997f, v)); }

/* an integral float clamped to [-1, 65536] and converted: bits above 15 are set exactly when
   the value lies outside [0, 65535]; inside, the low 16 bits are the u16 value */
__device__ __forceinline__ int clamp_q(float v) { return (int)__builtin_amdgcn_fmed3f(v, -1.0f, 65536.0f); }

template <typename T> struct Val;
template <> struct Val<uint16_t> {
    /* value as it sits in the u16 image after a float store */
    static __device__ __forceinline__ float store(float v) { return (float)f2u16(v); }
    /* round<T> (ComputeGradient.hpp:236-244); the range test is folded into acc, the OR of the
       clamped values (a bit above 15 set <=> some value left [0, 65535], where the reference
       throws) */
    static __device__ __forceinline__ float rnd(float v, int& acc) {
        v = round_away(v);
        acc |= clamp_q(v);
        return v;
    }
    static __device__ __forceinline__ uint16_t to_t(float v) { return f2u16(v); }
    static __device__ __forceinline__ bool failed(int acc) { return (acc & ~0xFFFF) != 0; }
};
template <> struct Val<float> {
    static __device__ __forceinline__ float store(float v) { return v; }
    static __device__ __forceinline__ float rnd(float v, int&) { return v; }
    static __device__ __forceinline__ float to_t(float v) { return v; }
    static __device__ __forceinline__ bool failed(int) { return false; }
};

/* input sample: u16 image + bspline offset (copyFromMeshWithUnaryOp, PixelData.hpp:788-811) */
template <typename T>
__device__ __forceinline__ float load_in(const T* p, float off, bool has_off) {
    if (has_off) return (float)f2u16((float)(*p) + off);
    return (float)(*p);
}

struct MinMaxInit {
    struct Args {
        uint32_t* mm;
    };
    static constexpr int THREADS = 64;
    static const char* name() { return "minmax_init"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t) {
        if (threadIdx.x == 0) {
            a.mm[0] = 65535u;
            a.mm[1] = 0u;
        }
    }
};

/* grid-stride over a virtual grid of G blocks (tile = block) */
__device__ __forceinline__ void minmax_body(const uint16_t* img, size_t n, uint32_t* mm, int64_t G, int64_t tile) {
    uint32_t lo = 65535u, hi = 0u;
    /* 16-byte loads over the aligned body, scalar tail */
    const size_t nv = n / 8;
    const uint4* v4 = (const uint4*)img;
    const size_t stride = (size_t)G * blockDim.x;
    size_t i = (size_t)tile * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < nv; i += 4 * stride) {
        uint4 a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = v4[i + u * stride];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t w[4] = {a[u].x, a[u].y, a[u].z, a[u].w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t l = w[q] & 0xFFFFu, h = w[q] >> 16;
                lo = min(lo, min(l, h));
                hi = max(hi, max(l, h));
            }
        }
    }
    for (; i < nv; i += stride) {
        const uint4 a = v4[i];
        const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t l = w[q] & 0xFFFFu, h = w[q] >> 16;
            lo = min(lo, min(l, h));
            hi = max(hi, max(l, h));
        }
    }
    for (size_t j = nv * 8 + (size_t)tile * blockDim.x + threadIdx.x; j < n; j += stride) {
        uint32_t v = img[j];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

struct MinMax {
    struct Args {
        const uint16_t* img;
        size_t n;
        uint32_t* mm;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "minmax"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { minmax_body(a.img, a.n, a.mm, a.G, t); }
};

struct Offset {
    struct Args {
        const uint32_t* mm;
        float lambda;
        float* off;
    };
    static constexpr int THREADS = 64;
    static const char* name() { return "bspline_offset"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t) {
        if (threadIdx.x != 0) return;
        if (a.lambda <= 0.0f) {
            *a.off = 0.0f;
            return;
        }
        int mn = (int)a.mm[0], mx = (int)a.mm[1];
        float o = (float)((65535 - (mx - mn)) / 2 - mn);
        *a.off = fmaxf(fminf(o, 100.0f), -100.0f);
    }
};

/* offset speculation for the fused y pass: before it, the min/max words are reset, the offset
   is set to the value every image with max + min <= 65335 gets (100 for lambda > 0) and the
   re-run flag is cleared */
struct OffsetGuess {
    struct Args {
        uint32_t* mm;
        float lambda;
        float* off;
        int* rerun;
    };
    static constexpr int THREADS = 64;
    static const char* name() { return "bspline_offset_guess"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t) {
        if (threadIdx.x != 0) return;
        a.mm[0] = 65535u;
        a.mm[1] = 0u;
        *a.off = a.lambda > 0.0f ? 100.0f : 0.0f;
        *a.rerun = 0;
    }
};
/* after the y pass: the offset from the min/max it gathered (computeL, AutoParameters.hpp:185-196);
   when the guess was wrong the offset is replaced, the y pass's overflow word cleared and the
   re-run flag set */
struct OffsetVerify {
    struct Args {
        const uint32_t* mm;
        float lambda;
        float* off;
        int* rerun;
        int* err_y;
    };
    static constexpr int THREADS = 64;
    static const char* name() { return "bspline_offset_verify"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t) {
        if (threadIdx.x != 0) return;
        float o = 0.0f;
        if (a.lambda > 0.0f) {
            const int mn = (int)a.mm[0], mx = (int)a.mm[1];
            o = fmaxf(fminf((float)((65535 - (mx - mn)) / 2 - mn), 100.0f), -100.0f);
        }
        if (o != *a.off) {
            *a.off = o;
            *a.err_y = 0;
            *a.rerun = 1;
        }
    }
};

struct OffsetCopy {
    struct Args {
        const uint16_t* in;
        uint16_t* out;
        size_t n;
        const float* d_off;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "offset_copy"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        const float off = *a.d_off;
        for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < a.n; i += (size_t)a.G * blockDim.x)
            a.out[i] = f2u16((float)a.in[i] + off);
    }
};

constexpr int TY = 64;   /* samples per tile row  */
constexpr int TL = 64;   /* lines per tile (wave) */

template <typename T>
__device__ __forceinline__ void iir_y_body(const T* __restrict__ in, T* out, int64_t y_num, int64_t n_lines, const IIRCoeffs& c,
                                           const float* d_off, int* d_err, int64_t blk) {
    __shared__ float tile[TL][TY + 1];
    const int lane = threadIdx.x;
    const int64_t line0 = blk * TL;
    const int64_t me = line0 + lane;
    const bool valid = me < n_lines;
    const bool has_off = d_off != nullptr;
    const float off = has_off ? *d_off : 0.0f;
    const float b1 = c.b1, b2 = c.b2, nf = c.nf;
    int err = 0;

    /* boundary sums over the first / last minLen samples (ComputeGradient.hpp:342-356) */
    float t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    if (valid) {
        const T* l = in + me * y_num;
        for (int k = 0; k < c.mL; ++k) {
            float v = load_in(l + k, off, has_off);
            t1 += c.bc1[k] * v;
            t2 += c.bc2[k] * v;
        }
        for (int k = 0; k < c.mL; ++k) {
            float v = load_in(l + (y_num - 1 - k), off, has_off);
            t3 += c.bc3[k] * v;
            t4 += c.bc4[k] * v;
        }
    }
    __syncthreads();

    /* causal pass (ComputeGradient.hpp:359-373): state unrounded */
    float s1 = t1, s2 = t2;
    const int64_t nch = (y_num + TY - 1) / TY;
    for (int64_t ch = 0; ch < nch; ++ch) {
        const int64_t y0 = ch * TY;
        const int ny = (int)min((int64_t)TY, y_num - y0);
        for (int l = 0; l < TL; ++l) {
            const int64_t ln = line0 + l;
            if (ln < n_lines && lane < ny) tile[l][lane] = load_in(in + ln * y_num + y0 + lane, off, has_off);
        }
        __syncthreads();
        if (valid) {
            for (int k = 0; k < ny; ++k) {
                const int64_t yy = y0 + k;
                float o;
                if (yy == 0) {
                    o = Val<T>::store(t2);
                } else if (yy == 1) {
                    o = Val<T>::store(t1);
                } else {
                    float t = s1 * b1 + s2 * b2 + tile[lane][k];
                    o = Val<T>::rnd(t, err);
                    s2 = s1;
                    s1 = t;
                }
                tile[lane][k] = o;
            }
        }
        __syncthreads();
        for (int l = 0; l < TL; ++l) {
            const int64_t ln = line0 + l;
            if (ln < n_lines && lane < ny) out[ln * y_num + y0 + lane] = Val<T>::to_t(tile[l][lane]);
        }
        __syncthreads();
    }

    /* end values (ComputeGradient.hpp:375-376) */
    const float e2 = Val<T>::rnd(t3 * nf, err);
    const float e1 = Val<T>::rnd(t4 * nf, err);
    const float bv_m2 = Val<T>::store(e2), bv_m1 = Val<T>::store(e1);

    /* anti-causal pass (ComputeGradient.hpp:382-407): state from stored/norm */
    float r2 = bv_m1 / nf;
    float r1 = bv_m2 / nf;
    for (int64_t ch = nch - 1; ch >= 0; --ch) {
        const int64_t y0 = ch * TY;
        const int ny = (int)min((int64_t)TY, y_num - y0);
        for (int l = 0; l < TL; ++l) {
            const int64_t ln = line0 + l;
            if (ln < n_lines && lane < ny) tile[l][lane] = (float)out[ln * y_num + y0 + lane];
        }
        __syncthreads();
        if (valid) {
            for (int k = ny - 1; k >= 0; --k) {
                const int64_t yy = y0 + k;
                if (yy >= y_num - 2) {
                    tile[lane][k] = (yy == y_num - 1) ? bv_m1 : bv_m2;
                } else {
                    float t = r1 * b1 + r2 * b2 + tile[lane][k];
                    tile[lane][k] = Val<T>::rnd(t * nf, err);
                    r2 = r1;
                    r1 = t;
                }
            }
        }
        __syncthreads();
        for (int l = 0; l < TL; ++l) {
            const int64_t ln = line0 + l;
            if (ln < n_lines && lane < ny) out[ln * y_num + y0 + lane] = Val<T>::to_t(tile[l][lane]);
        }
        __syncthreads();
    }
    if (Val<T>::failed(err)) atomicAdd(d_err, 1);
}

/* x / z pass: lines along an outer axis of length n and element stride s;
   thread = one y of one line (ComputeGradient.hpp:415-599, 601-779) */
template <typename T>
__device__ __forceinline__ void iir_outer_body(T* img, int64_t y_num, int64_t n, int64_t s, int64_t outer_stride, const IIRCoeffs& c,
                                               int* d_err, int64_t bx, int64_t by) {
    const int64_t y = bx * blockDim.x + threadIdx.x;
    if (y >= y_num) return;
    T* base = img + by * outer_stride + y;
    const float b1 = c.b1, b2 = c.b2, nf = c.nf;
    int err = 0;
    float v1 = 0, v2 = 0, v3 = 0, v4 = 0;
    for (int j = 0; j < c.mL; ++j) {
        const float a = (float)base[(int64_t)j * s];
        const float b = (float)base[(n - 1 - j) * s];
        v1 += c.bc1[j] * a;
        v2 += c.bc2[j] * a;
        v3 += c.bc3[j] * b;
        v4 += c.bc4[j] * b;
    }
    /* causal: y(0), y(1) from the boundary sums, then the state is the rounded output */
    base[0] = Val<T>::to_t(Val<T>::rnd(v2, err));
    base[s] = Val<T>::to_t(Val<T>::rnd(v1, err));
    float p1 = v1, p2 = v2;
    for (int64_t j = 2; j < n; ++j) {
        const float x = (float)base[j * s];
        const float t = Val<T>::rnd(x + b1 * p1 + b2 * p2, err);
        base[j * s] = Val<T>::to_t(t);
        p2 = p1;
        p1 = t;
    }
    /* anti-causal: state from the unrounded boundary sums */
    base[(n - 1) * s] = Val<T>::to_t(Val<T>::rnd(v4 * nf, err));
    base[(n - 2) * s] = Val<T>::to_t(Val<T>::rnd(v3 * nf, err));
    float q3 = v3, q4 = v4;
    for (int64_t j = n - 3; j >= 0; --j) {
        const float x = (float)base[j * s];
        const float t = (x + b1 * q3) + b2 * q4;
        base[j * s] = Val<T>::to_t(Val<T>::rnd(t * nf, err));
        q4 = q3;
        q3 = t;
    }
    if (Val<T>::failed(err)) atomicAdd(d_err, 1);
}


/* ---------------- vector paths ----------------
 * y pass: lane = one y-line, streamed with 16-byte loads/stores (8 u16 or 4 f32
 * samples); D vectors are loaded ahead of the recursion so several HBM
 * requests per lane stay in flight.  Requires y_num % VEC == 0 and 16-byte
 * aligned lines. */
template <typename T> struct Vec16;
template <> struct Vec16<uint16_t> {
    static constexpr int N = 8;
    static __device__ __forceinline__ float get(const uint4& v, int i) {
        const uint32_t w = (i < 2) ? v.x : (i < 4) ? v.y : (i < 6) ? v.z : v.w;
        return (float)((i & 1) ? (w >> 16) : (w & 0xFFFFu));
    }
    static __device__ __forceinline__ void set(uint4& v, int i, float f) {
        const uint32_t h = (uint32_t)f2u16(f);
        uint32_t* w = (i < 2) ? &v.x : (i < 4) ? &v.y : (i < 6) ? &v.z : &v.w;
        *w = (i & 1) ? ((*w & 0xFFFFu) | (h << 16)) : ((*w & 0xFFFF0000u) | h);
    }
    /* store of an integral value that is in [0, 65535] unless the pass reports an overflow */
    static __device__ __forceinline__ void set_r(uint4& v, int i, float f) {
        const uint32_t h = (uint32_t)clamp_q(f) & 0xFFFFu;
        uint32_t* w = (i < 2) ? &v.x : (i < 4) ? &v.y : (i < 6) ? &v.z : &v.w;
        *w = (i & 1) ? ((*w & 0xFFFFu) | (h << 16)) : ((*w & 0xFFFF0000u) | h);
    }
};
template <> struct Vec16<float> {
    static constexpr int N = 4;
    static __device__ __forceinline__ float get(const uint4& v, int i) {
        const uint32_t w = (i == 0) ? v.x : (i == 1) ? v.y : (i == 2) ? v.z : v.w;
        return __uint_as_float(w);
    }
    static __device__ __forceinline__ void set(uint4& v, int i, float f) {
        const uint32_t b = __float_as_uint(f);
        if (i == 0) v.x = b; else if (i == 1) v.y = b; else if (i == 2) v.z = b; else v.w = b;
    }
    static __device__ __forceinline__ void set_r(uint4& v, int i, float f) { set(v, i, f); }
};

template <typename T>
__device__ __forceinline__ float vin(const uint4& v, int i, float off, bool has_off) {
    const float x = Vec16<T>::get(v, i);
    /* x + offset lies in [-100, 65635]: the int32 conversion is the x86 one */
    return has_off ? (float)((uint32_t)(int)(x + off) & 0xFFFFu) : x;
}

template <typename T, int D>
__device__ __forceinline__ void iir_y_vec_body(const T* __restrict__ in, T* out, int64_t y_num, int64_t n_lines, const IIRCoeffs& c,
                                               const float* d_off, int* d_err, int64_t blk) {
    constexpr int V = Vec16<T>::N;
    const int64_t me = blk * blockDim.x + threadIdx.x;
    if (me >= n_lines) return;
    const bool has_off = d_off != nullptr;
    const float off = has_off ? *d_off : 0.0f;
    const float b1 = c.b1, b2 = c.b2, nf = c.nf;
    const uint4* src = (const uint4*)(in + me * y_num);
    uint4* dst = (uint4*)(out + me * y_num);
    const int64_t nv = y_num / V;
    int err = 0;
    /* boundary sums over the first / last minLen samples (ComputeGradient.hpp:342-356) */
    float t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    {
        const int nb = (c.mL + V - 1) / V;
        for (int b = 0; b < nb; ++b) {
            const uint4 v = src[b];
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const int k = b * V + i;
                if (k < c.mL) {
                    const float x = vin<T>(v, i, off, has_off);
                    t1 += c.bc1[k] * x;
                    t2 += c.bc2[k] * x;
                }
            }
        }
        for (int b = 0; b < nb; ++b) {
            const uint4 v = src[nv - 1 - b];
#pragma unroll
            for (int i = V - 1; i >= 0; --i) {
                const int k = b * V + (V - 1 - i);   /* distance from the end */
                if (k < c.mL) {
                    const float x = vin<T>(v, i, off, has_off);
                    t3 += c.bc3[k] * x;
                    t4 += c.bc4[k] * x;
                }
            }
        }
    }
    /* causal pass: state unrounded (ComputeGradient.hpp:359-373) */
    float s1 = t1, s2 = t2;
    /* y(0), y(1) are stored truncated, without the round<T> check (ComputeGradient.hpp:365-366) */
    const float t2q = Val<T>::store(t2), t1q = Val<T>::store(t1);
    for (int64_t b0 = 0; b0 < nv; b0 += D) {
        uint4 buf[D];
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (b0 + d < nv) buf[d] = src[b0 + d];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (b0 + d >= nv) break;
            uint4 o = buf[d];
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const int64_t k = (b0 + d) * V + i;
                float r;
                if (k == 0) {
                    r = t2q;
                } else if (k == 1) {
                    r = t1q;
                } else {
                    const float t = s1 * b1 + s2 * b2 + vin<T>(buf[d], i, off, has_off);
                    r = Val<T>::rnd(t, err);
                    s2 = s1;
                    s1 = t;
                }
                Vec16<T>::set_r(o, i, r);
            }
            dst[b0 + d] = o;
        }
    }
    /* end values (ComputeGradient.hpp:375-376) */
    const float e2 = Val<T>::rnd(t3 * nf, err);
    const float e1 = Val<T>::rnd(t4 * nf, err);
    const float bv_m2 = Val<T>::store(e2), bv_m1 = Val<T>::store(e1);
    /* anti-causal pass: state from stored/norm (ComputeGradient.hpp:382-407) */
    float r2 = bv_m1 / nf;
    float r1 = bv_m2 / nf;
    for (int64_t b0 = nv - 1; b0 >= 0; b0 -= D) {
        uint4 buf[D];
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (b0 - d >= 0) buf[d] = dst[b0 - d];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (b0 - d < 0) break;
            uint4 o = buf[d];
#pragma unroll
            for (int i = V - 1; i >= 0; --i) {
                const int64_t k = (b0 - d) * V + i;
                float r;
                if (k == y_num - 1) {
                    r = bv_m1;
                } else if (k == y_num - 2) {
                    r = bv_m2;
                } else {
                    const float t = r1 * b1 + r2 * b2 + Vec16<T>::get(buf[d], i);
                    r = Val<T>::rnd(t * nf, err);
                    r2 = r1;
                    r1 = t;
                }
                Vec16<T>::set_r(o, i, r);
            }
            dst[b0 - d] = o;
        }
    }
    if (Val<T>::failed(err)) atomicAdd(d_err, 1);
}

/* x / z pass, 16-byte vectors along y, D steps of the recursion loaded ahead */
template <typename T, int D>
__device__ __forceinline__ void iir_outer_vec_body(T* img, int64_t y_num, int64_t n, int64_t s, int64_t outer_stride, const IIRCoeffs& c,
                                                   int64_t lanes, int* d_err, int64_t t) {
    constexpr int V = Vec16<T>::N;
    const int64_t ny = y_num / V;
    /* lanes are numbered densely over (line, y vector), so short y rows still fill the blocks */
    const int64_t lane = t * blockDim.x + threadIdx.x;
    if (lane >= lanes) return;
    const int64_t by = lane / ny, yv = lane - by * ny;
    T* basep = img + by * outer_stride + yv * V;
    const int64_t sv = s / V;   /* stride in vectors */
    uint4* base = (uint4*)basep;
    const float b1 = c.b1, b2 = c.b2, nf = c.nf;
    int err = 0;
    float v1[V], v2[V], v3[V], v4[V];
#pragma unroll
    for (int i = 0; i < V; ++i) v1[i] = v2[i] = v3[i] = v4[i] = 0.0f;
    for (int j = 0; j < c.mL; ++j) {
        const uint4 a = base[(int64_t)j * sv];
        const uint4 b = base[(n - 1 - j) * sv];
        const float w1 = c.bc1[j], w2 = c.bc2[j], w3 = c.bc3[j], w4 = c.bc4[j];
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const float xa = Vec16<T>::get(a, i), xb = Vec16<T>::get(b, i);
            v1[i] += w1 * xa;
            v2[i] += w2 * xa;
            v3[i] += w3 * xb;
            v4[i] += w4 * xb;
        }
    }
    {
        uint4 o0 = make_uint4(0, 0, 0, 0), o1 = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < V; ++i) {
            Vec16<T>::set_r(o0, i, Val<T>::rnd(v2[i], err));
            Vec16<T>::set_r(o1, i, Val<T>::rnd(v1[i], err));
        }
        base[0] = o0;
        base[sv] = o1;
    }
    /* causal: state becomes the rounded output after the first step */
    float p1[V], p2[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        p1[i] = v1[i];
        p2[i] = v2[i];
    }
    for (int64_t j0 = 2; j0 < n; j0 += D) {
        uint4 buf[D];
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (j0 + d < n) buf[d] = base[(j0 + d) * sv];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (j0 + d >= n) break;
            uint4 o = make_uint4(0, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const float t = Val<T>::rnd(Vec16<T>::get(buf[d], i) + b1 * p1[i] + b2 * p2[i], err);
                Vec16<T>::set_r(o, i, t);
                p2[i] = p1[i];
                p1[i] = t;
            }
            base[(j0 + d) * sv] = o;
        }
    }
    /* anti-causal: state from the unrounded boundary sums */
    {
        uint4 oa = make_uint4(0, 0, 0, 0), ob = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < V; ++i) {
            Vec16<T>::set_r(oa, i, Val<T>::rnd(v4[i] * nf, err));
            Vec16<T>::set_r(ob, i, Val<T>::rnd(v3[i] * nf, err));
        }
        base[(n - 1) * sv] = oa;
        base[(n - 2) * sv] = ob;
    }
    for (int64_t j0 = n - 3; j0 >= 0; j0 -= D) {
        uint4 buf[D];
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (j0 - d >= 0) buf[d] = base[(j0 - d) * sv];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (j0 - d < 0) break;
            uint4 o = make_uint4(0, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const float t = (Vec16<T>::get(buf[d], i) + b1 * v3[i]) + b2 * v4[i];
                Vec16<T>::set_r(o, i, Val<T>::rnd(t * nf, err));
                v4[i] = v3[i];
                v3[i] = t;
            }
            base[(j0 - d) * sv] = o;
        }
    }
    if (Val<T>::failed(err)) atomicAdd(d_err, 1);
}

/* x / z pass without storing the causal result.  The causal recursion runs once over the
 * whole line and saves its state (the two previous outputs) at the start of every chunk of
 * C steps; the anti-causal sweep then walks the chunks from the end, re-reads each chunk's
 * input, recomputes its causal outputs from the saved state into registers and runs the
 * anti-causal recursion over them.  Per sample the line moves 2 reads and 1 write instead of
 * 2 reads and 2 writes, plus 4/C bytes of saved state.  Every value is computed by the same
 * operations in the same order as iir_outer_vec_body, so the outputs are identical; the causal
 * outputs held in registers are packed like the stored ones, and saved states are packed the
 * same way (exact for u16 unless the pass reports an overflow, where the reference throws).
 * In place: chunk k is written only after its input was read, and chunks below k are read
 * later but written later still. */
template <typename T, int C>
__device__ __forceinline__ void iir_outer_ck_body(T* img, int64_t y_num, int64_t n, int64_t s, int64_t outer_stride, const IIRCoeffs& c,
                                                  int* d_err, uint4* __restrict__ ck, int64_t ck_lanes, int64_t t) {
    constexpr int V = Vec16<T>::N;
    const int64_t ny = y_num / V;
    /* lanes are numbered densely over (line, y vector), so short y rows still fill the blocks */
    const int64_t lane = t * blockDim.x + threadIdx.x;
    if (lane >= ck_lanes) return;
    /* lane counts stay below 2^32 (the launcher checks): 32-bit division */
    const uint32_t by = (uint32_t)lane / (uint32_t)ny, yv = (uint32_t)lane - by * (uint32_t)ny;
    uint4* base = (uint4*)(img + (int64_t)by * outer_stride + (int64_t)yv * V);
    const int64_t sv = s / V;
    const float b1 = c.b1, b2 = c.b2, nf = c.nf;
    /* causal steps j = 2 .. n-1; the anti-causal sweep consumes j = 2 .. n-3 in chunks
       [2 + kC, min(2 + (k+1)C, n-2)), k < K.  Scratch per lane: the packed states at the
       starts of chunks 1 .. K-1, then the four boundary sums (unrounded floats), which are
       parked there between the sweeps to keep registers free for the chunk buffers. */
    const int64_t K = (n - 4 + C - 1) / C;
    uint4* sums = ck + (K - 1) * ck_lanes * 2 + lane * V;
    auto put_f = [&](uint4* dst, const float* v) {
#pragma unroll
        for (int q = 0; q < V / 4; ++q) dst[q] = make_uint4(__float_as_uint(v[4 * q]), __float_as_uint(v[4 * q + 1]),
                                                             __float_as_uint(v[4 * q + 2]), __float_as_uint(v[4 * q + 3]));
    };
    auto get_f = [&](const uint4* src, float* v) {
#pragma unroll
        for (int q = 0; q < V / 4; ++q) {
            const uint4 w = src[q];
            v[4 * q] = __uint_as_float(w.x);
            v[4 * q + 1] = __uint_as_float(w.y);
            v[4 * q + 2] = __uint_as_float(w.z);
            v[4 * q + 3] = __uint_as_float(w.w);
        }
    };
    int err = 0;
    float p1[V], p2[V];
    {
        float v1[V], v2[V], v3[V], v4[V];
#pragma unroll
        for (int i = 0; i < V; ++i) v1[i] = v2[i] = v3[i] = v4[i] = 0.0f;
        for (int j = 0; j < c.mL; ++j) {
            const uint4 a = base[(int64_t)j * sv];
            const uint4 b = base[(n - 1 - j) * sv];
            const float w1 = c.bc1[j], w2 = c.bc2[j], w3 = c.bc3[j], w4 = c.bc4[j];
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const float xa = Vec16<T>::get(a, i), xb = Vec16<T>::get(b, i);
                v1[i] += w1 * xa;
                v2[i] += w2 * xa;
                v3[i] += w3 * xb;
                v4[i] += w4 * xb;
            }
        }
        put_f(sums, v1);
        put_f(sums + V / 4, v2);
        put_f(sums + 2 * (V / 4), v3);
        put_f(sums + 3 * (V / 4), v4);
#pragma unroll
        for (int i = 0; i < V; ++i) {
            p1[i] = v1[i];
            p2[i] = v2[i];
        }
    }
    /* chunk loads are issued one chunk ahead of the recursion */
    auto load_chunk = [&](uint4* dst, int64_t j0, int cnt) {
#pragma unroll
        for (int d = 0; d < C; ++d)
            if (d < cnt) dst[d] = base[(j0 + d) * sv];
    };
    uint4 cur[C], nxt[C];
    {
        int cnt = (int)min((int64_t)C, n - 2);
        load_chunk(cur, 2, cnt);
        for (int64_t k = 0, j0 = 2; j0 < n; ++k, j0 += C) {
            const int64_t jn = j0 + C;
            const int cn = jn < n ? (int)min((int64_t)C, n - jn) : 0;
            load_chunk(nxt, jn, cn);
            if (k > 0 && k < K) {
                uint4 q1 = make_uint4(0, 0, 0, 0), q2 = make_uint4(0, 0, 0, 0);
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    Vec16<T>::set_r(q1, i, p1[i]);
                    Vec16<T>::set_r(q2, i, p2[i]);
                }
                uint4* cp = ck + ((k - 1) * ck_lanes + lane) * 2;
                cp[0] = q1;
                cp[1] = q2;
            }
#pragma unroll
            for (int d = 0; d < C; ++d) {
                if (d < cnt) {
#pragma unroll
                    for (int i = 0; i < V; ++i) {
                        const float t = Val<T>::rnd(Vec16<T>::get(cur[d], i) + b1 * p1[i] + b2 * p2[i], err);
                        p2[i] = p1[i];
                        p1[i] = t;
                    }
                }
            }
#pragma unroll
            for (int d = 0; d < C; ++d) cur[d] = nxt[d];
            cnt = cn;
        }
    }
    /* anti-causal: boundary values from the unrounded sums */
    float v3[V], v4[V];
    get_f(sums + 2 * (V / 4), v3);
    get_f(sums + 3 * (V / 4), v4);
    {
        uint4 oa = make_uint4(0, 0, 0, 0), ob = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < V; ++i) {
            Vec16<T>::set_r(oa, i, Val<T>::rnd(v4[i] * nf, err));
            Vec16<T>::set_r(ob, i, Val<T>::rnd(v3[i] * nf, err));
        }
        base[(n - 1) * sv] = oa;
        base[(n - 2) * sv] = ob;
    }
    auto back = [&](const uint4& f, int64_t j) {
        uint4 o = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const float t = (Vec16<T>::get(f, i) + b1 * v3[i]) + b2 * v4[i];
            Vec16<T>::set_r(o, i, Val<T>::rnd(t * nf, err));
            v4[i] = v3[i];
            v3[i] = t;
        }
        base[j * sv] = o;
    };
    /* two anti-causal steps (j, then j - 1) with the states swapping roles instead of moving */
    auto back2 = [&](const uint4& f, const uint4& g, int64_t j) {
        uint4 o = make_uint4(0, 0, 0, 0), u = o;
#pragma unroll
        for (int i = 0; i < V; ++i) {
            v4[i] = (Vec16<T>::get(f, i) + b1 * v3[i]) + b2 * v4[i];
            Vec16<T>::set_r(o, i, Val<T>::rnd(v4[i] * nf, err));
        }
        base[j * sv] = o;
#pragma unroll
        for (int i = 0; i < V; ++i) {
            v3[i] = (Vec16<T>::get(g, i) + b1 * v4[i]) + b2 * v3[i];
            Vec16<T>::set_r(u, i, Val<T>::rnd(v3[i] * nf, err));
        }
        base[(j - 1) * sv] = u;
    };
    /* causal outputs of chunk `buf` from state (p1, p2), then the anti-causal sweep over it */
    auto chunk = [&](uint4* buf, int64_t j0, int cnt) {
#pragma unroll
        for (int d = 0; d < C; ++d) {
            if (d < cnt) {
                uint4 o = make_uint4(0, 0, 0, 0);
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    const float t = Val<T>::rnd(Vec16<T>::get(buf[d], i) + b1 * p1[i] + b2 * p2[i], err);
                    Vec16<T>::set_r(o, i, t);
                    p2[i] = p1[i];
                    p1[i] = t;
                }
                buf[d] = o;
            }
        }
#pragma unroll
        for (int d = C - 1; d >= 0; --d)
            if (d < cnt) back(buf[d], j0 + d);
    };
    if (K > 0) {
        int64_t j0 = 2 + (K - 1) * C;
        int cnt = (int)min((int64_t)C, n - 2 - j0);
        load_chunk(cur, j0, cnt);
        uint4 q1 = make_uint4(0, 0, 0, 0), q2 = q1;
        if (K > 1) {
            const uint4* cp = ck + ((K - 2) * ck_lanes + lane) * 2;
            q1 = cp[0];
            q2 = cp[1];
        }
        /* chunks K-1 .. 1 start from packed states; chunk 0 from the boundary sums */
        for (int64_t k = K - 1; k >= 1; --k, j0 -= C) {
            uint4 n1 = make_uint4(0, 0, 0, 0), n2 = n1;
            load_chunk(nxt, j0 - C, C);
            if (k > 1) {
                const uint4* cp = ck + ((k - 2) * ck_lanes + lane) * 2;
                n1 = cp[0];
                n2 = cp[1];
            }
#pragma unroll
            for (int i = 0; i < V; ++i) {
                p1[i] = Vec16<T>::get(q1, i);
                p2[i] = Vec16<T>::get(q2, i);
            }
            chunk(cur, j0, cnt);
#pragma unroll
            for (int d = 0; d < C; ++d) cur[d] = nxt[d];
            q1 = n1;
            q2 = n2;
            cnt = C;
        }
        get_f(sums, p1);
        get_f(sums + V / 4, p2);
        chunk(cur, 2, cnt);
    }
    /* causal outputs at 0 and 1, as they would be stored */
    float v1[V], v2[V];
    get_f(sums, v1);
    get_f(sums + V / 4, v2);
    uint4 f0 = make_uint4(0, 0, 0, 0), f1 = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < V; ++i) {
        Vec16<T>::set_r(f0, i, Val<T>::rnd(v2[i], err));
        Vec16<T>::set_r(f1, i, Val<T>::rnd(v1[i], err));
    }
    if (n - 3 >= 1) back(f1, 1);
    back(f0, 0);
    if (Val<T>::failed(err)) atomicAdd(d_err, 1);
}

/* y pass without storing the causal result (u16, y_num % 64 == 0): the causal sweep streams
 * the lines once and saves each line's recursion state (two floats) at the start of every
 * chunk; the anti-causal sweep walks the chunks from the end, reloads each chunk of the input,
 * recomputes its causal outputs into the LDS row from the saved state and runs the
 * anti-causal recursion over the row before the chunk is written.  The same operations in the
 * same order as a plain causal-then-anti-causal sweep, so the outputs are identical; the line
 * moves 2 reads and one
 * write instead of 2 reads and 2 writes (plus 8 bytes of state per 64 samples).  In place is
 * safe: chunk ch is written after chunk ch - 1 has been loaded. */
__device__ __forceinline__ void iir_y_lds_ck_body(const uint16_t* in, uint16_t* out, int64_t y_num, int64_t n_lines, const IIRCoeffs& c,
                                                  const float* d_off, int* d_err, float2* __restrict__ ck, uint32_t* mm,
                                                  const int* run_if, uint16_t* pyr, int64_t px_num, int64_t pz_num, int64_t blk) {
    constexpr int CH = 64;
    constexpr int PITCH = 72;
    __shared__ __attribute__((aligned(16))) uint16_t tile[64 * PITCH];
    if (run_if && *run_if == 0) return;
    const int lane = threadIdx.x;
    /* LDS row R -> line (z * x_num + x), or -1.  Plain: 64 consecutive lines.  With pyr: rows
       0..31 are 32 consecutive x of plane 2 zp, rows 32..63 the same x of plane 2 zp + 1, so the
       2x2 (x, z) neighbourhoods of the mean pyramid's cells are all in the tile */
    const int64_t nxb = pyr ? (px_num + 31) / 32 : 1;
    const int64_t pzp = pyr ? blk / nxb : 0, pxb = pyr ? blk - pzp * nxb : 0;
    const int64_t line0 = blk * 64;
    auto line_of = [&](int R) -> int64_t {
        if (!pyr) return line0 + R < n_lines ? line0 + R : -1;
        const int64_t zz = 2 * pzp + (R >> 5), xx = pxb * 32 + (R & 31);
        return (zz < pz_num && xx < px_num) ? zz * px_num + xx : -1;
    };
    const int64_t me = line_of(lane);
    const bool valid = me >= 0;
    const bool has_off = d_off != nullptr;
    const float off = has_off ? *d_off : 0.0f;
    const float b1 = c.b1, b2 = c.b2, nf = c.nf;
    const int64_t nch = y_num / CH;
    const int seg = lane & 7;
    int err = 0;

    /* lines of the cooperative layout: load instruction i covers LDS rows 8i..8i+7 */
    int64_t lrow[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) lrow[i] = line_of(8 * i + (lane >> 3));
    auto gload = [&](const uint16_t* base, int64_t ch, uint4* r) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
            r[i] = (lrow[i] >= 0) ? *(const uint4*)(base + lrow[i] * y_num + ch * CH + seg * 8) : make_uint4(0, 0, 0, 0);
    };
    auto lds_put = [&](const uint4* r) {
#pragma unroll
        for (int i = 0; i < 8; ++i) *(uint4*)(&tile[(8 * i + (lane >> 3)) * PITCH + seg * 8]) = r[i];
    };
    auto gstore = [&](uint16_t* base, int64_t ch) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (lrow[i] >= 0) *(uint4*)(base + lrow[i] * y_num + ch * CH + seg * 8) = *(const uint4*)(&tile[(8 * i + (lane >> 3)) * PITCH + seg * 8]);
    };
    /* the mean pyramid's cells of one chunk (PixelData.hpp:944-962 with u16 storage): 16 x pairs
       x 32 y pairs of the plane pair; a lane sums 8 y pairs of one x pair over its 4 rows (eight
       u16 values per cell, exact in float) and stores them as one 16-byte vector */
    auto pyr_chunk = [&](int64_t ch) {
        const int xcl = lane & 15, ycg = lane >> 4;
        const int64_t Xc = (px_num + 1) / 2, Yc = y_num / 2;
        const int64_t xc = pxb * 16 + xcl;
        if (xc >= Xc) return;
        const int r00 = 2 * xcl;
        const int r01 = (pxb * 32 + 2 * xcl + 1 < px_num) ? r00 + 1 : r00;
        const bool hz = 2 * pzp + 1 < pz_num;
        const int rows[4] = {r00, r01, hz ? r00 + 32 : r00, hz ? r01 + 32 : r01};
        uint32_t sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint4 v = *(const uint4*)(&tile[rows[q] * PITCH + ycg * 16 + h * 8]);
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) sum[h * 4 + j] += (w[j] & 0xFFFFu) + (w[j] >> 16);
            }
        }
        uint32_t o[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = (uint32_t)(uint16_t)(int32_t)((float)sum[j] / 8.0f);
        *(uint4*)(pyr + (pzp * Xc + xc) * Yc + ch * (CH / 2) + ycg * 8) =
            make_uint4(o[0] | (o[1] << 16), o[2] | (o[3] << 16), o[4] | (o[5] << 16), o[6] | (o[7] << 16));
    };

    /* boundary sums (ComputeGradient.hpp:342-356) */
    float t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    if (valid) {
        const uint4* src = (const uint4*)(in + me * y_num);
        const int64_t nv = y_num / 8;
        const int nb = (c.mL + 7) / 8;
        for (int b = 0; b < nb; ++b) {
            const uint4 v = src[b];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = b * 8 + i;
                if (k < c.mL) {
                    const float x = vin<uint16_t>(v, i, off, has_off);
                    t1 += c.bc1[k] * x;
                    t2 += c.bc2[k] * x;
                }
            }
        }
        for (int b = 0; b < nb; ++b) {
            const uint4 v = src[nv - 1 - b];
#pragma unroll
            for (int i = 7; i >= 0; --i) {
                const int k = b * 8 + (7 - i);
                if (k < c.mL) {
                    const float x = vin<uint16_t>(v, i, off, has_off);
                    t3 += c.bc3[k] * x;
                    t4 += c.bc4[k] * x;
                }
            }
        }
    }
    const float t2q = Val<uint16_t>::store(t2), t1q = Val<uint16_t>::store(t1);
    /* causal recursion over one LDS row chunk from state (s1, s2); with `keep` the outputs
       replace the row's samples */
    /* the first chunk holds samples 0 and 1, which take the boundary values; the others run
       without per-sample tests.  Without `keep` (the first sweep) nothing is rounded: the
       recursion carries the unrounded values, and the range check of the rounded outputs
       becomes a running min / max of the unrounded ones, since roundf(t) lies in [0, 65535]
       exactly when -0.5 < t < 65535.5.  The re-computing sweep rounds the same values again
       and skips the check. */
    auto causal = [&](uint16_t* row, int64_t ch, float& s1, float& s2, bool keep, float& lo, float& hi) {
        int e = 0;
        if (ch == 0) {
#pragma unroll
            for (int q = 0; q < CH / 8; ++q) {
                uint4 v = *(uint4*)(row + q * 8);
                uint4 o = make_uint4(0, 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int k = q * 8 + i;
                    float r;
                    if (k == 0) {
                        r = t2q;
                    } else if (k == 1) {
                        r = t1q;
                    } else {
                        const float t = s1 * b1 + s2 * b2 + vin<uint16_t>(v, i, off, has_off);
                        if (keep) {
                            r = Val<uint16_t>::rnd(t, e);
                        } else {
                            lo = fminf(lo, t);
                            hi = fmaxf(hi, t);
                            r = 0.0f;
                        }
                        s2 = s1;
                        s1 = t;
                    }
                    if (keep) Vec16<uint16_t>::set_r(o, i, r);
                }
                if (keep) *(uint4*)(row + q * 8) = o;
            }
        } else {
#pragma unroll
            for (int q = 0; q < CH / 8; ++q) {
                uint4 v = *(uint4*)(row + q * 8);
                uint4 o = make_uint4(0, 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float t = s1 * b1 + s2 * b2 + vin<uint16_t>(v, i, off, has_off);
                    if (keep) {
                        Vec16<uint16_t>::set_r(o, i, Val<uint16_t>::rnd(t, e));
                    } else {
                        lo = fminf(lo, t);
                        hi = fmaxf(hi, t);
                    }
                    s2 = s1;
                    s1 = t;
                }
                if (keep) *(uint4*)(row + q * 8) = o;
            }
        }
    };

    /* causal sweep: states only; with mm, also the (min, max) of the raw samples (packed u16
       min / max over the loaded vectors; lines past the end load zeros, so they are skipped) */
    {
        float s1 = t1, s2 = t2, tlo = 0.0f, thi = 0.0f;
        uint4 nxt[8];
        typedef unsigned short us2 __attribute__((ext_vector_type(2)));
        us2 lo = {65535, 65535}, hi = {0, 0};
        gload(in, 0, nxt);
        for (int64_t ch = 0; ch < nch; ++ch) {
            __syncthreads();
            if (mm) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if (lrow[i] < 0) continue;
                    const uint32_t w[4] = {nxt[i].x, nxt[i].y, nxt[i].z, nxt[i].w};
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        us2 v;
                        __builtin_memcpy(&v, &w[q], 4);
                        lo = __builtin_elementwise_min(lo, v);
                        hi = __builtin_elementwise_max(hi, v);
                    }
                }
            }
            lds_put(nxt);
            if (ch + 1 < nch) gload(in, ch + 1, nxt);
            __syncthreads();
            if (pyr) pyr_chunk(ch);
            if (valid) {
                if (ch > 0) ck[(ch - 1) * n_lines + me] = make_float2(s1, s2);
                causal(&tile[lane * PITCH], ch, s1, s2, false, tlo, thi);
            }
        }
        if (!(tlo > -0.5f && thi < 65535.5f)) err |= 0x10000;
        if (mm) {
            uint32_t l = min((uint32_t)lo.x, (uint32_t)lo.y), h = max((uint32_t)hi.x, (uint32_t)hi.y);
            for (int o = 32; o > 0; o >>= 1) {
                l = min(l, (uint32_t)__shfl_xor((int)l, o));
                h = max(h, (uint32_t)__shfl_xor((int)h, o));
            }
            if (lane == 0) {
                atomicMin(&mm[0], l);
                atomicMax(&mm[1], h);
            }
        }
    }
    const float e2 = Val<uint16_t>::rnd(t3 * nf, err);
    const float e1 = Val<uint16_t>::rnd(t4 * nf, err);
    const float bv_m2 = Val<uint16_t>::store(e2), bv_m1 = Val<uint16_t>::store(e1);
    float r2 = bv_m1 / nf;
    float r1 = bv_m2 / nf;
    /* anti-causal sweep: recompute each chunk's causal outputs, then run backwards over them */
    uint4 nxt[8];
    gload(in, nch - 1, nxt);
    for (int64_t ch = nch - 1; ch >= 0; --ch) {
        __syncthreads();
        lds_put(nxt);
        if (ch - 1 >= 0) gload(in, ch - 1, nxt);
        __syncthreads();
        if (valid) {
            uint16_t* row = &tile[lane * PITCH];
            float s1 = t1, s2 = t2;
            if (ch > 0) {
                const float2 st = ck[(ch - 1) * n_lines + me];
                s1 = st.x;
                s2 = st.y;
            }
            float unused_lo = 0.0f, unused_hi = 0.0f;
            causal(row, ch, s1, s2, true, unused_lo, unused_hi);
            if (ch == nch - 1) {
                /* the last chunk holds samples y_num-2 and y_num-1, which take the boundary values */
#pragma unroll
                for (int q = CH / 8 - 1; q >= 0; --q) {
                    uint4 v = *(uint4*)(row + q * 8);
                    uint4 o = make_uint4(0, 0, 0, 0);
#pragma unroll
                    for (int i = 7; i >= 0; --i) {
                        const int k = q * 8 + i;
                        float r;
                        if (k == CH - 1) {
                            r = bv_m1;
                        } else if (k == CH - 2) {
                            r = bv_m2;
                        } else {
                            const float t = r1 * b1 + r2 * b2 + Vec16<uint16_t>::get(v, i);
                            r = Val<uint16_t>::rnd(t * nf, err);
                            r2 = r1;
                            r1 = t;
                        }
                        Vec16<uint16_t>::set_r(o, i, r);
                    }
                    *(uint4*)(row + q * 8) = o;
                }
            } else {
#pragma unroll
                for (int q = CH / 8 - 1; q >= 0; --q) {
                    uint4 v = *(uint4*)(row + q * 8);
                    uint4 o = make_uint4(0, 0, 0, 0);
#pragma unroll
                    for (int i = 7; i >= 0; --i) {
                        const float t = r1 * b1 + r2 * b2 + Vec16<uint16_t>::get(v, i);
                        Vec16<uint16_t>::set_r(o, i, Val<uint16_t>::rnd(t * nf, err));
                        r2 = r1;
                        r1 = t;
                    }
                    *(uint4*)(row + q * 8) = o;
                }
            }
        }
        __syncthreads();
        gstore(out, ch);
    }
    if (Val<uint16_t>::failed(err)) atomicAdd(d_err, 1);
}

/* per-volume arguments of the smoothing passes */
template <typename T>
struct IirArgs {
    const T* in;          /* y pass input (may equal out) */
    T* out;               /* y pass output / image of the x, z passes */
    int64_t y_num, n;     /* y length; length of the filtered axis (x, z passes) or lines (y) */
    int64_t s, outer_stride, gx;
    IIRCoeffs c;
    const float* d_off;
    int* d_err;
    uint4* ck = nullptr;      /* saved causal states (checkpointed x, z passes) */
    int64_t ck_lanes = 0;     /* lanes of a vector x / z launch; states of chunk k start at k * ck_lanes */
    uint32_t* mm = nullptr;   /* y pass: (min, max) of the raw input gathered on the way */
    const int* run_if = nullptr;   /* y pass: runs only when *run_if != 0 */
    uint16_t* pyr = nullptr;  /* y pass: also write the 2x2x2 mean level of the raw input here */
    int64_t px = 0, pz = 0;   /* y pass with pyr: x and z sizes (lines = px * pz) */
};

template <typename T>
struct IirY {
    using Args = IirArgs<T>;
    static constexpr int THREADS = 64;
    static const char* name() { return "bspline_y"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { iir_y_body<T>(a.in, a.out, a.y_num, a.n, a.c, a.d_off, a.d_err, t); }
};
template <typename T, int D>
struct IirYVec {
    using Args = IirArgs<T>;
    static constexpr int THREADS = 256;
    static const char* name() { return "bspline_y_vec"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { iir_y_vec_body<T, D>(a.in, a.out, a.y_num, a.n, a.c, a.d_off, a.d_err, t); }
};
struct IirYLdsCk {
    using Args = IirArgs<uint16_t>;
    static constexpr int THREADS = 64;
    static const char* name() { return "bspline_y_lds_ck"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        iir_y_lds_ck_body(a.in, a.out, a.y_num, a.n, a.c, a.d_off, a.d_err, (float2*)a.ck, a.mm, a.run_if, a.pyr, a.px, a.pz, t);
    }
};
template <typename T>
struct IirOuter {
    using Args = IirArgs<T>;
    static constexpr int THREADS = 256;
    static const char* name() { return "bspline_outer"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        iir_outer_body<T>(a.out, a.y_num, a.n, a.s, a.outer_stride, a.c, a.d_err, t % a.gx, t / a.gx);
    }
};
template <typename T, int D>
struct IirOuterVec {
    using Args = IirArgs<T>;
    static constexpr int THREADS = 256;
    static const char* name() { return "bspline_outer_vec"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        iir_outer_vec_body<T, D>(a.out, a.y_num, a.n, a.s, a.outer_stride, a.c, a.ck_lanes, a.d_err, t);
    }
};

template <typename T, int C>
struct IirOuterCk {
    using Args = IirArgs<T>;
    static constexpr int THREADS = 256;
    static constexpr int MIN_WAVES = 2;   /* two chunk buffers in registers; one wave per SIMD hides too little latency */
    static const char* name() { return "bspline_outer_ck"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        iir_outer_ck_body<T, C>(a.out, a.y_num, a.n, a.s, a.outer_stride, a.c, a.d_err, a.ck, a.ck_lanes, t);
    }
};

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

void launch_minmax_offset(aprh_context* ctx, const uint16_t* img, size_t n, float lambda, float* d_offset, uint32_t* d_mm) {
    /* initial (min, max) set by a kernel, so the sequence can be captured in a graph */
    launch_tiles<MinMaxInit>(ctx, {d_mm}, 1);
    const int64_t G = stride_tiles((int64_t)n, 256 * 8 * 4, 2048);
    launch_tiles<MinMax>(ctx, {img, n, d_mm, G}, G);
    launch_tiles<Offset>(ctx, {d_mm, lambda, d_offset}, 1);
}

void launch_offset_copy(aprh_context* ctx, const uint16_t* in, uint16_t* out, size_t n, const float* d_offset) {
    const int64_t G = stride_tiles((int64_t)n, 256 * 4);
    launch_tiles<OffsetCopy>(ctx, {in, out, n, d_offset, G}, G);
}

bool iir_y_fuses_offset(int64_t y, const void* in, const void* out) { return y % 64 == 0 && aligned16(in) && aligned16(out); }

void launch_iir_y_offset(aprh_context* ctx, const uint16_t* in, uint16_t* out, int64_t y, int64_t x, int64_t z, const IIRCoeffs& c,
                         float lambda, float* d_offset, uint32_t* d_mm, int* d_rerun, int* d_err, uint16_t* pyr) {
    const int64_t lines = x * z;
    IirArgs<uint16_t> a{in, out, y, lines, 0, 0, 1, c, d_offset, d_err};
    Buf ck(ctx->pool, (size_t)std::max<int64_t>(y / 64 - 1, 1) * (size_t)lines * 8);
    a.ck = ck.as<uint4>();
    launch_tiles<OffsetGuess>(ctx, {d_mm, lambda, d_offset, d_rerun}, 1);
    a.mm = d_mm;
    if (pyr) {
        /* blocks over (plane pair, 32 x) so the first level of the mean pyramid comes along */
        a.pyr = pyr;
        a.px = x;
        a.pz = z;
        launch_tiles<IirYLdsCk>(ctx, a, cdiv(z, 2) * cdiv(x, 32));
    } else {
        launch_tiles<IirYLdsCk>(ctx, a, cdiv(lines, 64));
    }
    launch_tiles<OffsetVerify>(ctx, {d_mm, lambda, d_offset, d_rerun, d_err}, 1);
    /* the pass again with the true offset, as blocks that return at once unless the guess was wrong */
    a.mm = nullptr;
    a.pyr = nullptr;
    a.run_if = d_rerun;
    launch_tiles<IirYLdsCk>(ctx, a, cdiv(lines, 64));
}

template <typename T>
void launch_iir_y(aprh_context* ctx, const T* in, T* out, int64_t y, int64_t x, int64_t z, const IIRCoeffs& c,
                  const float* d_offset, int* d_err) {
    const int64_t lines = x * z;
    constexpr int V = Vec16<T>::N;
    IirArgs<T> a{in, out, y, lines, 0, 0, 1, c, d_offset, d_err};
    if constexpr (std::is_same<T, uint16_t>::value) {
        if (y % 64 == 0 && aligned16(in) && aligned16(out)) {
            /* line states at the start of chunks 1 .. y/64-1 */
            Buf ck(ctx->pool, (size_t)std::max<int64_t>(y / 64 - 1, 1) * (size_t)lines * 8);
            a.ck = ck.as<uint4>();
            launch_tiles<IirYLdsCk>(ctx, a, cdiv(lines, 64));
            return;
        }
    }
    if (y % V == 0 && y >= 2 * V && aligned16(in) && aligned16(out)) {
        launch_tiles<IirYVec<T, 4>>(ctx, a, cdiv(lines, 256));
    } else {
        launch_tiles<IirY<T>>(ctx, a, cdiv(lines, TL));
    }
}

/* checkpointed x / z pass: chunk length and the shortest axis it is used for */
constexpr int OUTER_CK = 16;
constexpr int64_t OUTER_CK_MIN = 64;

template <typename T>
void launch_iir_outer(aprh_context* ctx, T* img, int64_t y, int64_t x, int64_t z, int axis, const IIRCoeffs& c, int* d_err) {
    constexpr int V = Vec16<T>::N;
    /* x pass: lines along x (stride y), one per (y, z); z pass: along z (stride x*y), one per (y, x) */
    const int64_t n = axis == 1 ? x : z, s = axis == 1 ? y : x * y, outer = axis == 1 ? z : x;
    const int64_t ostride = axis == 1 ? x * y : y;
    if (y % V == 0 && aligned16(img)) {
        const int64_t lanes = (y / V) * outer;
        IirArgs<T> a{img, img, y, n, s, ostride, 1, c, nullptr, d_err};
        a.ck_lanes = lanes;
        if (n >= OUTER_CK_MIN && lanes < (int64_t(1) << 32)) {
            /* a batch plan also times checkpoints every 8 and every 32 steps (the buffer then
               holds the states of the densest spacing) */
            const bool alts = std::is_same<T, uint16_t>::value && ctx->recorder;
            const int64_t chunks = (n - 4 + (alts ? 8 : OUTER_CK) - 1) / (alts ? 8 : OUTER_CK);
            /* the chunk states live until the launch ends: from the pool, returned after the
               launch is queued (stream order keeps later users behind it) */
            Buf ck(ctx->pool, ((size_t)(chunks - 1) * 2 + V) * (size_t)lanes * 16);
            a.ck = ck.as<uint4>();
            const int64_t tiles = cdiv(lanes, 256);
            if (alts)
                launch_tiles_alts(ctx,
                                  {tile_alt<IirOuterCk<T, 8>>(a, tiles, "checkpoint every 8"),
                                   tile_alt<IirOuterCk<T, OUTER_CK>>(a, tiles, "checkpoint every 16"),
                                   tile_alt<IirOuterCk<T, 32>>(a, tiles, "checkpoint every 32")},
                                  1);
            else
                launch_tiles<IirOuterCk<T, OUTER_CK>>(ctx, a, tiles);
        } else {
            launch_tiles<IirOuterVec<T, 8>>(ctx, a, cdiv(lanes, 256));
        }
        return;
    }
    const int64_t gx = cdiv(y, 256);
    launch_tiles<IirOuter<T>>(ctx, IirArgs<T>{img, img, y, n, s, ostride, gx, c, nullptr, d_err}, gx * outer);
}

template void launch_iir_y<uint16_t>(aprh_context*, const uint16_t*, uint16_t*, int64_t, int64_t, int64_t, const IIRCoeffs&,
                                     const float*, int*);
template void launch_iir_y<float>(aprh_context*, const float*, float*, int64_t, int64_t, int64_t, const IIRCoeffs&, const float*,
                                  int*);
template void launch_iir_outer<uint16_t>(aprh_context*, uint16_t*, int64_t, int64_t, int64_t, int, const IIRCoeffs&, int*);
template void launch_iir_outer<float>(aprh_context*, float*, int64_t, int64_t, int64_t, int, const IIRCoeffs&, int*);

}  // namespace aprh
