/*
 * k_iir.hip -- B-spline smoothing of the conversion pipeline on gfx950.
 *
 *  - min/max reduction + bspline offset     (AutoParameters.hpp:165-196)
 *  - y pass: one lane per y-line; 64 lines per wave are staged through an
 *    LDS tile (64 lines x 64 samples, padded rows) so global loads/stores are
 *    coalesced along y while each lane walks its own line sequentially
 *    (ComputeGradient.hpp:248-413).  The offset copy of computeL
 *    (APRConverter.hpp:213-224) is fused into the tile load.
 *  - x / z passes: one lane per (y, outer) line; lanes run along y, so every
 *    step of the recursion is a coalesced row access
 *    (ComputeGradient.hpp:415-599, 601-779).
 *
 * Every lane keeps the reference's per-line FP32 operation order; the build
 * disables FP contraction, so results are bit-identical to the CPU path.
 */
#include "aprh_internal.hpp"

namespace aprh {

namespace {

__device__ __forceinline__ uint16_t f2u16(float v) {
    /* g++/x86-64 float->uint16 conversion: truncation through int32 */
    if (!(v > -2147483648.0f && v < 2147483648.0f)) return 0;
    return (uint16_t)(int32_t)v;
}

template <typename T> struct Val;
template <> struct Val<uint16_t> {
    /* value as it sits in the u16 image after a float store */
    static __device__ __forceinline__ float store(float v) { return (float)f2u16(v); }
    /* round<T> (ComputeGradient.hpp:236-244) */
    static __device__ __forceinline__ float rnd(float v, int& err) {
        v = roundf(v);
        if (v < 0.0f || v > 65535.0f) err++;
        return v;
    }
    static __device__ __forceinline__ uint16_t to_t(float v) { return f2u16(v); }
};
template <> struct Val<float> {
    static __device__ __forceinline__ float store(float v) { return v; }
    static __device__ __forceinline__ float rnd(float v, int&) { return v; }
    static __device__ __forceinline__ float to_t(float v) { return v; }
};

/* input sample: u16 image + bspline offset (copyFromMeshWithUnaryOp, PixelData.hpp:788-811) */
template <typename T>
__device__ __forceinline__ float load_in(const T* p, float off, bool has_off) {
    if (has_off) return (float)f2u16((float)(*p) + off);
    return (float)(*p);
}

__global__ void k_minmax(const uint16_t* img, size_t n, uint32_t* mm) {
    uint32_t lo = 65535u, hi = 0u;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t v = img[i];
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

__global__ void k_offset(const uint32_t* mm, float lambda, float* off) {
    if (lambda <= 0.0f) {
        *off = 0.0f;
        return;
    }
    int mn = (int)mm[0], mx = (int)mm[1];
    float o = (float)((65535 - (mx - mn)) / 2 - mn);
    *off = fmaxf(fminf(o, 100.0f), -100.0f);
}

__global__ void k_offset_copy(const uint16_t* in, uint16_t* out, size_t n, const float* d_off) {
    const float off = *d_off;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = f2u16((float)in[i] + off);
}

constexpr int TY = 64;   /* samples per tile row  */
constexpr int TL = 64;   /* lines per tile (wave) */

template <typename T>
__global__ __launch_bounds__(64) void k_iir_y(const T* __restrict__ in, T* out, int64_t y_num, int64_t n_lines, IIRCoeffs c,
                                              const float* d_off, int* d_err) {
    __shared__ float tile[TL][TY + 1];
    const int lane = threadIdx.x;
    const int64_t line0 = (int64_t)blockIdx.x * TL;
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
    if (err) atomicAdd(d_err, err);
}

/* x / z pass: lines along an outer axis of length n and element stride s;
   thread = one y of one line (ComputeGradient.hpp:415-599, 601-779) */
template <typename T>
__global__ __launch_bounds__(256) void k_iir_outer(T* img, int64_t y_num, int64_t n, int64_t s, int64_t outer_stride,
                                                   IIRCoeffs c, int* d_err) {
    const int64_t y = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= y_num) return;
    T* base = img + (int64_t)blockIdx.y * outer_stride + y;
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
    if (err) atomicAdd(d_err, err);
}

}  // namespace

void launch_minmax_offset(aprh_context* ctx, const uint16_t* img, size_t n, float lambda, float* d_offset, uint32_t* d_mm) {
    const uint32_t init[2] = {65535u, 0u};
    HIP_CHECK(hipMemcpyAsync(d_mm, init, 8, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_minmax, dim3(2048), dim3(256), 0, ctx->stream, img, n, d_mm);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_offset, dim3(1), dim3(1), 0, ctx->stream, d_mm, lambda, d_offset);
    HIP_LAUNCH_CHECK();
}

void launch_offset_copy(aprh_context* ctx, const uint16_t* in, uint16_t* out, size_t n, const float* d_offset) {
    hipLaunchKernelGGL(k_offset_copy, dim3(4096), dim3(256), 0, ctx->stream, in, out, n, d_offset);
    HIP_LAUNCH_CHECK();
}

template <typename T>
void launch_iir_y(aprh_context* ctx, const T* in, T* out, int64_t y, int64_t x, int64_t z, const IIRCoeffs& c,
                  const float* d_offset, int* d_err) {
    const int64_t lines = x * z;
    hipLaunchKernelGGL(k_iir_y<T>, dim3((unsigned)cdiv(lines, TL)), dim3(64), 0, ctx->stream, in, out, y, lines, c, d_offset, d_err);
    HIP_LAUNCH_CHECK();
}

template <typename T>
void launch_iir_outer(aprh_context* ctx, T* img, int64_t y, int64_t x, int64_t z, int axis, const IIRCoeffs& c, int* d_err) {
    if (axis == 1) {
        /* x pass: outer = z */
        hipLaunchKernelGGL(k_iir_outer<T>, dim3((unsigned)cdiv(y, 256), (unsigned)z), dim3(256), 0, ctx->stream, img, y, x, y, x * y,
                           c, d_err);
    } else {
        /* z pass: outer = x */
        hipLaunchKernelGGL(k_iir_outer<T>, dim3((unsigned)cdiv(y, 256), (unsigned)x), dim3(256), 0, ctx->stream, img, y, z, x * y, y,
                           c, d_err);
    }
    HIP_LAUNCH_CHECK();
}

template void launch_iir_y<uint16_t>(aprh_context*, const uint16_t*, uint16_t*, int64_t, int64_t, int64_t, const IIRCoeffs&,
                                     const float*, int*);
template void launch_iir_y<float>(aprh_context*, const float*, float*, int64_t, int64_t, int64_t, const IIRCoeffs&, const float*,
                                  int*);
template void launch_iir_outer<uint16_t>(aprh_context*, uint16_t*, int64_t, int64_t, int64_t, int, const IIRCoeffs&, int*);
template void launch_iir_outer<float>(aprh_context*, float*, int64_t, int64_t, int64_t, int, const IIRCoeffs&, int*);

}  // namespace aprh
