/*
 * k_smooth.hip -- APRReconstruction::reconstruct_smooth (APRReconstruction.hpp:270-295).
 *
 * The piecewise-constant image is smoothed along y, x and z by running-sum means whose
 * half-width at a pixel is off = min(T(floor(2^(l_max - level) / scale)), T(offset_max))
 * (level = the level of the particle covering the pixel, offset_max = min(20, n/2 - 1))
 * -- calc_sat_adaptive_y / _x / _z (:528-876).
 *
 * Each pass is two kernels.  The cumulative sums of a line are a sequential float
 * recursion, one lane per line (the reference's own order); they go to a float
 * scratch laid out like the image.  The reference keeps a ring of 2 offset_max + 2
 * cumulative rows for x and z; its entries are the same prefix sums of the pass's
 * input (a ring slot is read only before it is reused), so the full prefix array
 * gives the same values.  The output of a pixel then depends only on those sums,
 * its own offset and its position, and a second, element-wise kernel applies the
 * reference's sequence of stores for that pixel, converting to the image type T at
 * every store as the reference's assignments do (float / double -> uint16 through
 * int32 truncation, the x86-64 conversion).
 */
#include "aprh_internal.hpp"

namespace aprh {

namespace {

__device__ __forceinline__ uint16_t sm_f2u16(float v) {
    if (!(v > -2147483648.0f && v < 2147483648.0f)) return 0;
    return (uint16_t)(int32_t)v;
}
__device__ __forceinline__ uint16_t sm_d2u16(double v) {
    if (!(v > -2147483648.0 && v < 2147483648.0)) return 0;
    return (uint16_t)(int32_t)v;
}

template <typename T> struct SmT;
template <> struct SmT<float> {
    static __device__ __forceinline__ float f(float v) { return v; }
    static __device__ __forceinline__ float d(double v) { return (float)v; }
};
template <> struct SmT<uint16_t> {
    static __device__ __forceinline__ uint16_t f(float v) { return sm_f2u16(v); }
    static __device__ __forceinline__ uint16_t d(double v) { return sm_d2u16(v); }
};

/* std::min((T)floor(pow(2, d_max - level) / scale), (T)offset_max) */
template <typename T>
__device__ __forceinline__ T sm_offset(uint16_t level, float scale, int om, int d_max) {
    const T a = SmT<T>::d(floor(ldexp(1.0, d_max - (int)level) / (double)scale));
    const T b = (T)om;
    return a < b ? a : b;
}

constexpr int SL = 64;   /* lines (lanes) per block of the y prefix kernel, and y samples per tile */

/* prefix sums along y, one lane per line; 64 x 64 tiles staged through LDS so that loads and
   stores run along y */
template <typename T>
__global__ __launch_bounds__(SL) void k_sm_prefix_y(const T* __restrict__ m, float* __restrict__ tv, int64_t yn, int64_t nlines) {
    __shared__ float tile[SL][SL + 1];
    const int lane = threadIdx.x;
    const int64_t line0 = (int64_t)blockIdx.x * SL;
    const int nl = (int)min((int64_t)SL, nlines - line0);
    float t = 0.0f;
    for (int64_t y0 = 0; y0 < yn; y0 += SL) {
        const int ny = (int)min((int64_t)SL, yn - y0);
        for (int l = 0; l < nl; ++l)
            if (lane < ny) tile[l][lane] = (float)m[(line0 + l) * yn + y0 + lane];
        __syncthreads();
        if (lane < nl)
            for (int k = 0; k < ny; ++k) {
                t += tile[lane][k];
                tile[lane][k] = t;
            }
        __syncthreads();
        for (int l = 0; l < nl; ++l)
            if (lane < ny) tv[(line0 + l) * yn + y0 + lane] = tile[l][lane];
        __syncthreads();
    }
}

/* calc_sat_adaptive_y (APRReconstruction.hpp:528-646): the stores of pixel k in the reference's
   order, from the line's prefix sums */
template <typename T>
__global__ __launch_bounds__(256) void k_sm_out_y(T* __restrict__ m, const float* __restrict__ tv, const uint16_t* __restrict__ lev,
                                                  int64_t yn, int64_t n, float scale, int om, int d_max) {
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = idx % yn;
        const float* tl = tv + (idx - k);
        const float off = (float)sm_offset<T>(lev[idx], scale, om, d_max);
        const float div = 2 * off + 1;
        const float kf = (float)k;
        T s = m[idx];
        if (k == 0) s = (T)0;
        if (k >= 1 && k <= om + 1 && kf <= off + 1) s = SmT<T>::f(-tl[0] / div);
        if (kf >= off + 1) s = SmT<T>::f(-tl[(int64_t)(kf - off - 1)] / div);
        if (kf < (float)yn - off) s = SmT<T>::f((float)s + tl[(int64_t)(kf + off)] / div);
        if (k >= yn - om && kf >= (float)yn - off) {
            const float counter = kf - ((float)yn - off) + 1;
            s = SmT<T>::f((float)s * div);
            s = SmT<T>::f((float)s + tl[yn - 1]);
            s = SmT<T>::d((double)s * (1.0 / (double)(div - counter)));
        }
        if (k >= 1 && k < om + 1 && kf < off + 1) s = SmT<T>::d((double)s * ((double)div / (1.0 * k + (double)off)));
        if (k == 0) s = SmT<T>::f((float)s * (div / (off + 1)));
        m[idx] = s;
    }
}

/* prefix sums along x (axis 1) or z (axis 2): lane = (outer line, y), sequential over the axis;
   consecutive lanes are consecutive y */
template <typename T>
__global__ __launch_bounds__(256) void k_sm_prefix_outer(const T* __restrict__ m, float* __restrict__ tv, int64_t yn, int64_t xn,
                                                         int64_t zn, int axis) {
    const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t outer = axis == 1 ? zn : xn;
    if (lane >= outer * yn) return;
    const int64_t k = lane % yn, o = lane / yn;
    const int64_t n = axis == 1 ? xn : zn;
    const int64_t stride = axis == 1 ? yn : xn * yn;
    const int64_t base = (axis == 1 ? o * xn * yn : o * yn) + k;
    float t = 0.0f;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t p = base + i * stride;
        t = (i == 0) ? (float)m[p] : (float)m[p] + t;
        tv[p] = t;
    }
}

/* calc_sat_adaptive_x / _z (APRReconstruction.hpp:649-876): low end, middle and high end stores of
   pixel i, the later ones overriding; the z pass tests the x index at its low end, as the
   reference does (:821) */
template <typename T>
__global__ __launch_bounds__(256) void k_sm_out_outer(T* __restrict__ m, const float* __restrict__ tv, const uint16_t* __restrict__ lev,
                                                      int64_t yn, int64_t xn, int64_t zn, int axis, float scale, int om, int d_max) {
    const int64_t N = yn * xn * zn;
    const int64_t n = axis == 1 ? xn : zn;
    const int64_t stride = axis == 1 ? yn : xn * yn;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < N; idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = idx / yn;
        const int64_t x = r % xn, z = r / xn;
        const int64_t i = axis == 1 ? x : z;
        const float* tl = tv + (idx - i * stride);   /* the line's prefix sums, element j at tl[j * stride] */
        const int64_t off = (int64_t)sm_offset<T>(lev[idx], scale, om, d_max);
        T s = m[idx];
        const int64_t lo_index = axis == 1 ? i : x;
        if (i <= om && lo_index < off + 1) s = SmT<T>::f(tl[(i + off) * stride] / (float)(i + off + 1));
        if (i >= 1 && i >= off + 1 && i < n - off)
            s = SmT<T>::f((tl[(i + off) * stride] - tl[(i - off - 1) * stride]) / (float)(2 * off + 1));
        if (i >= n - om && i >= n - off) s = SmT<T>::f((tl[(n - 1) * stride] - tl[(i - off - 1) * stride]) / (float)(n - i + off));
        m[idx] = s;
    }
}

template <typename T>
void smooth_passes(aprh_context* ctx, T* img, const uint16_t* lev, float* tv, int64_t yn, int64_t xn, int64_t zn, const float* scale3,
                   int offset_max, int d_max) {
    hipStream_t st = ctx->stream;
    const int64_t N = yn * xn * zn;
    const unsigned g = (unsigned)std::min<int64_t>(cdiv(N, 256), 16384);
    if (yn > 1) {
        const int om = std::min<int>(offset_max, (int)(yn / 2 - 1));
        const int64_t lines = xn * zn;
        hipLaunchKernelGGL(k_sm_prefix_y<T>, dim3((unsigned)cdiv(lines, SL)), dim3(SL), 0, st, img, tv, yn, lines);
        HIP_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sm_out_y<T>, dim3(g), dim3(256), 0, st, img, tv, lev, yn, N, scale3[0], om, d_max);
        HIP_LAUNCH_CHECK();
    }
    for (int axis = 1; axis <= 2; ++axis) {
        const int64_t n = axis == 1 ? xn : zn;
        if (n <= 1) continue;
        const int om = std::min<int>(offset_max, (int)(n / 2 - 1));
        const int64_t lanes = (axis == 1 ? zn : xn) * yn;
        hipLaunchKernelGGL(k_sm_prefix_outer<T>, dim3((unsigned)cdiv(lanes, 256)), dim3(256), 0, st, img, tv, yn, xn, zn, axis);
        HIP_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sm_out_outer<T>, dim3(g), dim3(256), 0, st, img, tv, lev, yn, xn, zn, axis, scale3[axis], om, d_max);
        HIP_LAUNCH_CHECK();
    }
}

}  // namespace

/* APRReconstruction::reconstruct_smooth: reconstruct_constant, reconstruct_level, then the
   adaptive means along y, x and z (offset_max 20, d_max = l_max) */
void reconstruct_smooth(aprh_context* ctx, aprh_apr& a, const void* parts, int pdt, const float* scale3, void* img, int idt) {
    const int L = a.l_max;
    const int64_t yn = a.y_num[L], xn = a.x_num[L], zn = a.z_num[L];
    const int64_t N = yn * xn * zn;
    if (N == 0) return;
    if (cdiv(xn * zn, SL) >= (int64_t(1) << 31) || xn * yn >= (int64_t(1) << 40))
        throw Error(APRH_ERR_UNSUPPORTED, "image too large for reconstruct_smooth");
    reconstruct_any(ctx, a, parts, pdt, L, img, idt);
    Buf lev(ctx->pool, (size_t)N * 2), tv(ctx->pool, (size_t)N * 4);
    reconstruct_level(ctx, a, L, lev.ptr(), APRH_DTYPE_U16);
    if (idt == APRH_DTYPE_U16)
        smooth_passes<uint16_t>(ctx, (uint16_t*)img, lev.as<uint16_t>(), tv.as<float>(), yn, xn, zn, scale3, 20, L);
    else
        smooth_passes<float>(ctx, (float*)img, lev.as<uint16_t>(), tv.as<float>(), yn, xn, zn, scale3, 20, L);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace aprh
