/*
 * apr_oracle.hpp -- CPU restatement of LibAPR's APR hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker / CPU baseline.  The product (libapr_amd) never
 * includes, links or calls it.
 *
 * Every function restates one reference function and cites the reference
 * file:line it follows.  Floating-point expressions keep the reference's
 * operand order and its C++ type promotions (float vs double sub-expressions,
 * float->integer truncations) so that, compiled with -ffp-contract=off and
 * without -ffast-math, results are bit-identical to the reference built the
 * same way.  Float->u16 / float->u32 conversions reproduce what g++ emits on
 * x86-64 (truncation through a 32-bit resp. 64-bit signed integer).
 *
 * Pinning of each part (checked in tests/): see the header of
 * oracle_capi.cpp and the tests that call each entry point.
 */
#ifndef APR_ORACLE_HPP
#define APR_ORACLE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/aprh_params.h"

namespace oracle {

/* ------------------------------------------------------------------------ */
/* Dense y-fastest volume, idx = z*x_num*y_num + x*y_num + y                  */
/* (PixelData.hpp:509,521)                                                    */
/* ------------------------------------------------------------------------ */
template <typename T>
struct Mesh {
    int64_t y_num = 0, x_num = 0, z_num = 0;
    std::vector<T> v;
    Mesh() {}
    Mesh(int64_t y, int64_t x, int64_t z, T init = T(0)) { resize(y, x, z, init); }
    void resize(int64_t y, int64_t x, int64_t z, T init = T(0)) {
        y_num = y; x_num = x; z_num = z;
        v.assign((size_t)(y * x * z), init);
    }
    size_t size() const { return v.size(); }
    T& at(int64_t y, int64_t x, int64_t z) { return v[(size_t)(z * x_num * y_num + x * y_num + y)]; }
    const T& at(int64_t y, int64_t x, int64_t z) const { return v[(size_t)(z * x_num * y_num + x * y_num + y)]; }
};

/* x86-64 g++ conversion semantics for out-of-contract values. */
static inline uint16_t f2u16(float v) {
    if (!(v > -2147483648.0f && v < 2147483648.0f)) return 0;
    return (uint16_t)(int32_t)v;
}
static inline uint16_t d2u16(double v) {
    if (!(v > -2147483648.0 && v < 2147483648.0)) return 0;
    return (uint16_t)(int32_t)v;
}
static inline uint32_t f2u32(float v) {
    if (!(v > -9.2233720368547758e18f && v < 9.2233720368547758e18f)) return 0;
    return (uint32_t)(int64_t)v;
}

template <typename T> struct Store;
template <> struct Store<uint16_t> {
    static uint16_t cvt(float v) { return f2u16(v); }
    /* ComputeGradient.hpp:226-244: std::round then range check, value kept as float */
    static float rnd(float v, size_t& err) {
        v = std::round(v);
        if (v < 0.0f || v > 65535.0f) err++;
        return v;
    }
};
template <> struct Store<float> {
    static float cvt(float v) { return v; }
    static float rnd(float v, size_t&) { return v; }
};

/* ------------------------------------------------------------------------ */
/* B-spline smoothing coefficients (ComputeGradient.hpp:256-330).             */
/* The reference calls the global double-precision libm functions on float   */
/* operands; each step below states that promotion explicitly.               */
/* ------------------------------------------------------------------------ */
struct BsplineCoeffs {
    float b1 = 0, b2 = 0, norm_factor = 0;
    size_t k0 = 0, minLen = 0;
    std::vector<float> bc1, bc2, bc3, bc4;
};

static inline float impulse_resp_f(float k, float rho, float omg) {
    /* ComputeGradient.hpp:213-217 */
    float ak = std::fabs(k);
    double num = std::pow((double)rho, (double)ak) * std::sin((double)((ak + 1.0f) * omg));
    return (float)(num / std::sin((double)omg));
}

static inline float impulse_resp_back_f(float k, float rho, float omg, float gamma, float c0) {
    /* ComputeGradient.hpp:219-223 */
    float ak = std::fabs(k);
    double p = std::pow((double)rho, (double)ak);
    double t1 = (double)c0 * p;
    double t2 = std::cos((double)(omg * ak)) + (double)gamma * std::sin((double)(omg * ak));
    double r2 = std::pow((double)rho, 2.0);
    double den = std::pow(1.0 - 2.0 * (double)rho * std::cos((double)omg) + r2, 2.0);
    return (float)((t1 * t2) * (1.0 / den));
}

static inline BsplineCoeffs bspline_coeffs(float lambda, float tol, size_t len) {
    BsplineCoeffs c;
    /* eq 4.6 / 4.5 (ComputeGradient.hpp:256-258) */
    float f96 = 96.0f * lambda;
    float f24 = 24.0f * lambda;
    float f144p3 = 3.0f + 144.0f * lambda;
    float xi = (float)((double)(1.0f - f96) + (double)f24 * std::sqrt((double)f144p3));
    float inv_xi = 1.0f / xi;
    double d1 = (double)(f24 - 1.0f) - std::sqrt((double)xi);
    double d2 = d1 / (double)f24;
    double d4 = (double)(48.0f * lambda) + (double)f24 * std::sqrt((double)f144p3);
    float rho = (float)(d2 * std::sqrt((double)inv_xi * d4));
    float omg = (float)std::atan(std::sqrt((double)(inv_xi * (144.0f * lambda - 1.0f))));
    /* eq 4.8 (ComputeGradient.hpp:260-261) */
    double pr2 = std::pow((double)rho, 2.0);
    double co = std::cos((double)omg);
    float r2f = 2.0f * rho;
    double c0d = (((1.0 + pr2) / (1.0 - pr2)) * (1.0 - (double)r2f * co + pr2)) / (1.0 + (double)r2f * co + pr2);
    float c0 = (float)c0d;
    float gamma = (float)(((1.0 - pr2) / (1.0 + pr2)) * (1.0 / std::tan((double)omg)));
    /* ComputeGradient.hpp:263-264 */
    c.b1 = (float)((double)r2f * co);
    c.b2 = (float)(-std::pow((double)rho, 2.0));
    /* ComputeGradient.hpp:271-274 */
    size_t k0 = (size_t)std::ceil(std::fabs(std::log((double)tol) / std::log((double)rho)));
    c.k0 = k0;
    c.minLen = std::min(k0, len);
    c.norm_factor = (float)std::pow(1.0 - 2.0 * (double)rho * co + pr2, 2.0);

    std::vector<float> f(k0 + 3), b(k0 + 3);
    for (size_t k = 0; k < k0 + 3; ++k) f[k] = impulse_resp_f((float)k, rho, omg);
    for (size_t k = 0; k < k0 + 3; ++k) b[k] = impulse_resp_back_f((float)k, rho, omg, gamma, c0);
    const size_t mL = c.minLen;
    /* ComputeGradient.hpp:287-330 */
    c.bc1.assign(k0, 0.0f);
    c.bc1[1] = f[0];
    for (size_t k = 0; k < k0; ++k) c.bc1[k] += f[k + 1];
    for (size_t k = mL; k < k0; ++k) c.bc1[mL - 1] += c.bc1[k];
    c.bc2.assign(k0, 0.0f);
    for (size_t k = 0; k < k0; ++k) c.bc2[k] = f[k];
    for (size_t k = mL; k < k0; ++k) c.bc2[mL - 1] += c.bc2[k];
    c.bc3.assign(k0, 0.0f);
    c.bc3[0] = b[1];
    for (size_t k = 0; k < k0 - 1; ++k) c.bc3[k + 1] += b[k] + b[k + 2];
    for (size_t k = mL; k < k0; ++k) c.bc3[mL - 1] += c.bc3[k];
    c.bc4.assign(k0, 0.0f);
    c.bc4[0] = b[0];
    for (size_t k = 1; k < k0; ++k) c.bc4[k] += 2.0f * b[k];
    for (size_t k = mL; k < k0; ++k) c.bc4[mL - 1] += c.bc4[k];
    return c;
}

/* ------------------------------------------------------------------------ */
/* Recursive smoothing-spline filters                                        */
/* ------------------------------------------------------------------------ */

/* bspline_filt_rec_y (ComputeGradient.hpp:248-413). Causal state kept      */
/* unrounded, outputs rounded; anti-causal state restarts from stored/norm.  */
template <typename T>
void bspline_filt_rec_y(Mesh<T>& img, float lambda, float tol, size_t& err) {
    const int64_t z_num = img.z_num, x_num = img.x_num, y_num = img.y_num;
    BsplineCoeffs c = bspline_coeffs(lambda, tol, (size_t)y_num);
    const float b1 = c.b1, b2 = c.b2, nf = c.norm_factor;
    const int64_t mL = (int64_t)c.minLen;
    size_t e = 0;
#pragma omp parallel for reduction(+ : e) schedule(static)
    for (int64_t z = 0; z < z_num; ++z) {
        for (int64_t x = 0; x < x_num; ++x) {
            T* line = img.v.data() + z * x_num * y_num + x * y_num;
            float t1 = 0, t2 = 0, t3 = 0, t4 = 0;
            for (int64_t k = 0; k < mL; ++k) {
                t1 += c.bc1[k] * (float)line[k];
                t2 += c.bc2[k] * (float)line[k];
            }
            for (int64_t k = 0; k < mL; ++k) {
                t3 += c.bc3[k] * (float)line[y_num - 1 - k];
                t4 += c.bc4[k] * (float)line[y_num - 1 - k];
            }
            line[0] = Store<T>::cvt(t2);
            line[1] = Store<T>::cvt(t1);
            for (int64_t k = 2; k < y_num; ++k) {
                float t = t1 * b1 + t2 * b2 + (float)line[k];
                line[k] = Store<T>::cvt(Store<T>::rnd(t, e));
                t2 = t1;
                t1 = t;
            }
            line[y_num - 2] = Store<T>::cvt(Store<T>::rnd(t3 * nf, e));
            line[y_num - 1] = Store<T>::cvt(Store<T>::rnd(t4 * nf, e));
        }
    }
#pragma omp parallel for reduction(+ : e) schedule(static)
    for (int64_t z = 0; z < z_num; ++z) {
        for (int64_t x = 0; x < x_num; ++x) {
            T* line = img.v.data() + z * x_num * y_num + x * y_num;
            float t2 = (float)line[y_num - 1] / nf;
            float t1 = (float)line[y_num - 2] / nf;
            for (int64_t k = y_num - 3; k >= 0; --k) {
                float t = t1 * b1 + t2 * b2 + (float)line[k];
                line[k] = Store<T>::cvt(Store<T>::rnd(t * nf, e));
                t2 = t1;
                t1 = t;
            }
        }
    }
    err += e;
}

/* Shared body of bspline_filt_rec_x (ComputeGradient.hpp:601-779) and       */
/* bspline_filt_rec_z (:415-599): lines along an outer axis of length n with  */
/* element stride s; each line is a vector over y. The causal recursion keeps */
/* the *rounded* output as state from the third sample on.                   */
template <typename T>
static void bspline_outer_axis(T* base, int64_t n, int64_t s, int64_t y_num, const BsplineCoeffs& c,
                               std::vector<float>& v1, std::vector<float>& v2, std::vector<float>& v3,
                               std::vector<float>& v4, size_t& e) {
    const float b1 = c.b1, b2 = c.b2, nf = c.norm_factor;
    const int64_t mL = (int64_t)c.minLen;
    std::fill(v1.begin(), v1.end(), 0.0f);
    std::fill(v2.begin(), v2.end(), 0.0f);
    std::fill(v3.begin(), v3.end(), 0.0f);
    std::fill(v4.begin(), v4.end(), 0.0f);
    for (int64_t j = 0; j < mL; ++j) {
        const T* a = base + j * s;
        const T* bk = base + (n - 1 - j) * s;
        for (int64_t k = 0; k < y_num; ++k) {
            v1[k] += c.bc1[j] * (float)a[k];
            v2[k] += c.bc2[j] * (float)a[k];
            v3[k] += c.bc3[j] * (float)bk[k];
            v4[k] += c.bc4[j] * (float)bk[k];
        }
    }
    for (int64_t k = 0; k < y_num; ++k) base[k] = Store<T>::cvt(Store<T>::rnd(v2[k], e));
    for (int64_t k = 0; k < y_num; ++k) base[s + k] = Store<T>::cvt(Store<T>::rnd(v1[k], e));
    for (int64_t j = 2; j < n; ++j) {
        T* row = base + j * s;
        for (int64_t k = 0; k < y_num; ++k) {
            v2[k] = Store<T>::rnd((float)row[k] + b1 * v1[k] + b2 * v2[k], e);
        }
        std::swap(v1, v2);
        for (int64_t k = 0; k < y_num; ++k) row[k] = Store<T>::cvt(v1[k]);
    }
    for (int64_t k = 0; k < y_num; ++k) base[(n - 1) * s + k] = Store<T>::cvt(Store<T>::rnd(v4[k] * nf, e));
    for (int64_t k = 0; k < y_num; ++k) base[(n - 2) * s + k] = Store<T>::cvt(Store<T>::rnd(v3[k] * nf, e));
    for (int64_t j = n - 3; j >= 0; --j) {
        T* row = base + j * s;
        for (int64_t k = 0; k < y_num; ++k) {
            float t = ((float)row[k] + b1 * v3[k]) + b2 * v4[k];
            row[k] = Store<T>::cvt(Store<T>::rnd(t * nf, e));
            v4[k] = v3[k];
            v3[k] = t;
        }
    }
}

template <typename T>
void bspline_filt_rec_x(Mesh<T>& img, float lambda, float tol, size_t& err) {
    const int64_t z_num = img.z_num, x_num = img.x_num, y_num = img.y_num;
    BsplineCoeffs c = bspline_coeffs(lambda, tol, (size_t)x_num);
    size_t e = 0;
#pragma omp parallel reduction(+ : e)
    {
        std::vector<float> v1(y_num), v2(y_num), v3(y_num), v4(y_num);
#pragma omp for schedule(static)
        for (int64_t z = 0; z < z_num; ++z)
            bspline_outer_axis(img.v.data() + z * x_num * y_num, x_num, y_num, y_num, c, v1, v2, v3, v4, e);
    }
    err += e;
}

template <typename T>
void bspline_filt_rec_z(Mesh<T>& img, float lambda, float tol, size_t& err) {
    const int64_t z_num = img.z_num, x_num = img.x_num, y_num = img.y_num;
    BsplineCoeffs c = bspline_coeffs(lambda, tol, (size_t)z_num);
    size_t e = 0;
#pragma omp parallel reduction(+ : e)
    {
        std::vector<float> v1(y_num), v2(y_num), v3(y_num), v4(y_num);
#pragma omp for schedule(static)
        for (int64_t x = 0; x < x_num; ++x)
            bspline_outer_axis(img.v.data() + x * y_num, z_num, x_num * y_num, y_num, c, v1, v2, v3, v4, e);
    }
    err += e;
}

/* get_smooth_bspline_3D (ComputeGradient.hpp:173-205) + overflow exceptions */
/* (ComputeGradient.hpp:409-412, :595-598, :775-778).                        */
template <typename T>
void smooth_bspline_3d(Mesh<T>& img, float lambda) {
    const float tol = 0.0001f;
    size_t err = 0;
    if (img.y_num > 2) {
        bspline_filt_rec_y(img, lambda, tol, err);
        if (err) throw std::invalid_argument("integer under-/overflow encountered in bspline_filt_rec_y");
    }
    if (img.x_num > 2) {
        bspline_filt_rec_x(img, lambda, tol, err);
        if (err) throw std::invalid_argument("integer under-/overflow encountered in bspline_filt_rec_x");
    }
    if (img.z_num > 2) {
        bspline_filt_rec_z(img, lambda, tol, err);
        if (err) throw std::invalid_argument("integer under-/overflow encountered in bspline_filt_rec_z");
    }
}

/* ------------------------------------------------------------------------ */
/* Inverse cubic B-spline [1,4,1]/6 (ComputeGradient.hpp:784-964)            */
/* ------------------------------------------------------------------------ */
static inline void inv_bspline_y(Mesh<float>& m) {
    const int64_t z_num = m.z_num, x_num = m.x_num, y_num = m.y_num;
    const float a1 = (float)(1.0 / 6.0), a2 = (float)(4.0 / 6.0), a3 = (float)(1.0 / 6.0);
#pragma omp parallel
    {
        std::vector<float> t(y_num);
#pragma omp for schedule(static)
        for (int64_t z = 0; z < z_num; ++z) {
            for (int64_t x = 0; x < x_num; ++x) {
                float* l = m.v.data() + z * x_num * y_num + x * y_num;
                for (int64_t k = 0; k < y_num; ++k) t[k] = l[k];
                l[0] = a2 * t[0];
                l[0] += (a1 + a3) * t[1];
                for (int64_t k = 1; k < y_num - 1; ++k) l[k] = a1 * t[k - 1] + a2 * t[k] + a3 * t[k + 1];
                l[y_num - 1] = (a1 + a3) * t[y_num - 2];
                l[y_num - 1] += a2 * t[y_num - 1];
            }
        }
    }
}

/* x (ComputeGradient.hpp:898-964) and z (:828-895) share one form:          */
/* out[0] = a1*in[1] + a2*in[0] + a3*in[1]; out[j] = a1*in[j-1]+a2*in[j]+a3*in[j+1];
   out[n-1] = (a1+a3)*in[n-2] + a2*in[n-1].                                  */
static inline void inv_bspline_outer(float* base, int64_t n, int64_t s, int64_t y_num, std::vector<float>& t1,
                                     std::vector<float>& t2, std::vector<float>& t3) {
    const float a1 = (float)(1.0 / 6.0), a2 = (float)(4.0 / 6.0), a3 = (float)(1.0 / 6.0);
    for (int64_t k = 0; k < y_num; ++k) {
        t1[k] = base[s + k];
        t2[k] = base[k];
    }
    for (int64_t j = 0; j < n - 1; ++j) {
        for (int64_t k = 0; k < y_num; ++k) t3[k] = base[(j + 1) * s + k];
        for (int64_t k = 0; k < y_num; ++k) base[j * s + k] = a1 * t1[k] + a2 * t2[k] + a3 * t3[k];
        for (int64_t k = 0; k < y_num; ++k) {
            t1[k] = t2[k];
            t2[k] = t3[k];
        }
    }
    for (int64_t k = 0; k < y_num; ++k) {
        float v = (a1 + a3) * t1[k];
        v += a2 * t2[k];
        base[(n - 1) * s + k] = v;
    }
}

static inline void inv_bspline_x(Mesh<float>& m) {
    const int64_t z_num = m.z_num, x_num = m.x_num, y_num = m.y_num;
#pragma omp parallel
    {
        std::vector<float> t1(y_num), t2(y_num), t3(y_num);
#pragma omp for schedule(static)
        for (int64_t z = 0; z < z_num; ++z) inv_bspline_outer(m.v.data() + z * x_num * y_num, x_num, y_num, y_num, t1, t2, t3);
    }
}

static inline void inv_bspline_z(Mesh<float>& m) {
    const int64_t z_num = m.z_num, x_num = m.x_num, y_num = m.y_num;
#pragma omp parallel
    {
        std::vector<float> t1(y_num), t2(y_num), t3(y_num);
#pragma omp for schedule(static)
        for (int64_t x = 0; x < x_num; ++x) inv_bspline_outer(m.v.data() + x * y_num, z_num, x_num * y_num, y_num, t1, t2, t3);
    }
}

/* ------------------------------------------------------------------------ */
/* Gradient magnitude + 2x2x2 max downsample (ComputeGradient.hpp:975-1051). */
/* Central differences in float, squared/summed/sqrt in double, truncated.   */
/* ------------------------------------------------------------------------ */
template <typename S>
static inline S grad_store(double v);
template <> inline uint16_t grad_store<uint16_t>(double v) { return d2u16(v); }
template <> inline float grad_store<float>(double v) { return (float)v; }

template <typename S>
void grad_mag_ds(const Mesh<S>& in, Mesh<S>& grad, float hx, float hy, float hz) {
    const int64_t z_num = in.z_num, x_num = in.x_num, y_num = in.y_num;
    const int64_t x_ds = grad.x_num, y_ds = grad.y_num, z_ds = grad.z_num;
    const int64_t xy = x_num * y_num;
    const float h2x = 2.0f * hx, h2y = 2.0f * hy, h2z = 2.0f * hz;
    /* (S - S) is int for uint16 and float for float; then / (2*h) in float */
    auto term = [](float d, float h2) { return std::pow((double)(d / h2), 2.0); };
    auto df = [](S a, S b) { return (float)(a - b); };
#pragma omp parallel
    {
        std::vector<S> tmp(y_num);
#pragma omp for schedule(static)
        for (int64_t zd = 0; zd < z_ds; ++zd) {
            for (int64_t z = 2 * zd; z <= std::min(2 * zd + 1, z_num - 1); ++z) {
                const int64_t zm = z > 0 ? z - 1 : 0, zp = std::min(z + 1, z_num - 1);
                for (int64_t x = 0; x < x_num; ++x) {
                    const S* up = in.v.data() + zm * xy + x * y_num;
                    const S* down = in.v.data() + zp * xy + x * y_num;
                    const S* left = in.v.data() + z * xy + (x > 0 ? x - 1 : 0) * y_num;
                    const S* right = in.v.data() + z * xy + std::min(x + 1, x_num - 1) * y_num;
                    const S* cen = in.v.data() + z * xy + x * y_num;
                    if (y_num >= 2) {
                        tmp[0] = grad_store<S>(std::sqrt(term(df(right[0], left[0]), h2x) +
                                                         term(df(down[0], up[0]), h2z) +
                                                         term(df(cen[1], cen[0]), h2y)));
                        int64_t e = y_num - 1;
                        tmp[e] = grad_store<S>(std::sqrt(term(df(right[e], left[e]), h2x) +
                                                         term(df(down[e], up[e]), h2z) +
                                                         term(df(cen[e], cen[e - 1]), h2y)));
                    } else {
                        tmp[0] = 0;
                    }
                    for (int64_t y = 1; y < y_num - 1; ++y) {
                        tmp[y] = grad_store<S>(std::sqrt(term(df(right[y], left[y]), h2x) +
                                                         term(df(down[y], up[y]), h2z) +
                                                         term(df(cen[y + 1], cen[y - 1]), h2y)));
                    }
                    const int64_t z2 = z / 2, x2 = x / 2;
                    for (int64_t k = 0; k < y_ds; ++k) {
                        const int64_t ks = std::min(2 * k + 1, y_num - 1);
                        S& g = grad.v[(size_t)(z2 * x_ds * y_ds + x2 * y_ds + k)];
                        g = std::max(tmp[2 * k], std::max(tmp[ks], g));
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* 2x2x2 downsample with fixed reduce order (PixelData.hpp:915-967)          */
/* ------------------------------------------------------------------------ */
template <typename T, typename S, typename R, typename C>
void downsample(const Mesh<T>& in, Mesh<S>& out, R reduce, C cop) {
    const int64_t z_num = in.z_num, x_num = in.x_num, y_num = in.y_num;
    const int64_t zd = (int64_t)std::ceil(z_num / 2.0), xd = (int64_t)std::ceil(x_num / 2.0), yd = (int64_t)std::ceil(y_num / 2.0);
    out.resize(yd, xd, zd);
    const int64_t xy = x_num * y_num;
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < zd; ++z) {
        for (int64_t x = 0; x < xd; ++x) {
            const int64_t shx = std::min(2 * x + 1, x_num - 1), shz = std::min(2 * z + 1, z_num - 1);
            for (int64_t y = 0; y < yd; ++y) {
                const int64_t shy = std::min(2 * y + 1, y_num - 1);
                const T* a = in.v.data();
                out.v[(size_t)(z * xd * yd + x * yd + y)] = cop(
                    reduce(reduce(reduce(reduce(reduce(reduce(reduce(a[2 * z * xy + 2 * x * y_num + 2 * y], a[2 * z * xy + 2 * x * y_num + shy]),
                                                              a[2 * z * xy + shx * y_num + 2 * y]),
                                                       a[2 * z * xy + shx * y_num + shy]),
                                                a[shz * xy + 2 * x * y_num + 2 * y]),
                                         a[shz * xy + 2 * x * y_num + shy]),
                                  a[shz * xy + shx * y_num + 2 * y]),
                           a[shz * xy + shx * y_num + shy]));
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Local intensity scale (LocalIntensityScale.hpp:26-494)                    */
/* ------------------------------------------------------------------------ */
struct LisWindow {
    float var_rescale;
    int win[6];
    bool active[3];
};

/* get_window_alt (LocalIntensityScale.hpp:250-303) */
static inline LisWindow lis_window(const aprh_params& par, int64_t y, int64_t x, int64_t z) {
    static const double r3[6] = {12.8214, 26.1256, 40.2795, 23.3692, 36.2061, 27.0385};
    static const double r2[6] = {13.2421, 28.7069, 52.0385, 24.4272, 34.9565, 21.1891};
    static const double r1[6] = {13.9040, 31.2843, 57.4037, 30.3767, 45.9930, 29.0890};
    static const int w1[] = {1, 1, 1, 2, 2, 3}, w2[] = {2, 3, 4, 4, 5, 6};
    int psf_ind = (int)std::max((float)(std::round((double)(par.psfx / par.dx)) - 1), 0.0f);
    psf_ind = std::min(psf_ind, 5);
    int wv = std::max(w1[psf_ind], w2[psf_ind]);
    LisWindow w;
    for (int i = 0; i < 6; ++i) w.win[i] = 0;
    int64_t dims[3] = {y, x, z};
    int nact = 0;
    for (int a = 0; a < 3; ++a) {
        w.active[a] = dims[a] > wv;
        if (w.active[a]) {
            w.win[a] = w1[psf_ind];
            w.win[a + 3] = w2[psf_ind];
            nact++;
        }
    }
    w.var_rescale = (float)(nact == 3 ? r3[psf_ind] : nact == 2 ? r2[psf_ind] : r1[psf_ind]);
    return w;
}

/* paddPixels (PixelData.hpp:1070-1176): reflect about the edge pixel */
template <typename T>
void padd(const Mesh<T>& in, Mesh<T>& out, int sy, int sx, int sz) {
    sy = in.y_num > 1 ? std::min<int64_t>(sy, in.y_num - 1) : 0;
    sx = in.x_num > 1 ? std::min<int64_t>(sx, in.x_num - 1) : 0;
    sz = in.z_num > 1 ? std::min<int64_t>(sz, in.z_num - 1) : 0;
    out.resize(in.y_num + 2 * sy, in.x_num + 2 * sx, in.z_num + 2 * sz);
    for (int64_t j = 0; j < in.z_num; ++j)
        for (int64_t i = 0; i < in.x_num; ++i)
            for (int64_t k = 0; k < in.y_num; ++k) out.at(k + sy, i + sx, j + sz) = in.at(k, i, j);
    if (in.y_num > 1) {
        for (int64_t j = 0; j < out.z_num; ++j)
            for (int64_t i = 0; i < out.x_num; ++i) {
                for (int64_t k = 0; k < sy; ++k) out.at(k, i, j) = out.at(2 * sy - k, i, j);
                int64_t idx = sy + 2;
                for (int64_t k = out.y_num - sy; k < out.y_num; ++k) out.at(k, i, j) = out.at(out.y_num - idx++, i, j);
            }
    }
    if (in.x_num > 1) {
        for (int64_t j = 0; j < out.z_num; ++j) {
            for (int64_t i = 0; i < sx; ++i)
                for (int64_t k = 0; k < out.y_num; ++k) out.at(k, i, j) = out.at(k, 2 * sx - i, j);
            int64_t idx = sx + 2;
            for (int64_t i = out.x_num - sx; i < out.x_num; ++i) {
                for (int64_t k = 0; k < out.y_num; ++k) out.at(k, i, j) = out.at(k, out.x_num - idx, j);
                idx++;
            }
        }
    }
    if (in.z_num > 1) {
        for (int64_t j = 0; j < sz; ++j)
            for (int64_t i = 0; i < out.x_num; ++i)
                for (int64_t k = 0; k < out.y_num; ++k) out.at(k, i, j) = out.at(k, i, 2 * sz - j);
        for (int64_t j = out.z_num - sz; j < out.z_num; ++j) {
            int64_t idx = sz + 2 + j - (out.z_num - sz);
            for (int64_t i = 0; i < out.x_num; ++i)
                for (int64_t k = 0; k < out.y_num; ++k) out.at(k, i, j) = out.at(k, i, out.z_num - idx);
        }
    }
}

/* calc_sat_mean_y (LocalIntensityScale.hpp:311-363) */
static inline void sat_mean_y(Mesh<float>& m, int64_t off) {
    const int64_t z_num = m.z_num, x_num = m.x_num, y_num = m.y_num;
    const float divisor = (float)(2 * off + 1);
#pragma omp parallel
    {
        std::vector<float> tv(y_num);
#pragma omp for schedule(static)
        for (int64_t j = 0; j < z_num; ++j) {
            for (int64_t i = 0; i < x_num; ++i) {
                float* l = m.v.data() + j * x_num * y_num + i * y_num;
                float t = 0;
                for (int64_t k = 0; k < y_num; ++k) {
                    t += l[k];
                    tv[k] = t;
                }
                for (int64_t k = 0; k <= off; ++k) l[k] = 0;
                for (int64_t k = off + 1; k < y_num; ++k) l[k] = -tv[k - off - 1] / divisor;
                for (int64_t k = 0; k < y_num - off; ++k) l[k] += tv[k + off] / divisor;
                float counter = 0;
                for (int64_t k = y_num - off; k < y_num; ++k) {
                    counter++;
                    l[k] *= divisor;
                    l[k] += tv[y_num - 1];
                    l[k] = (float)((double)l[k] * (1.0 / (double)(divisor - counter)));
                }
                for (int64_t k = 1; k <= off; ++k) l[k] = (float)((double)l[k] * ((double)divisor / ((double)(k + off) + 1.0)));
                l[0] = (float)((double)l[0] * ((double)divisor / ((double)off + 1.0)));
            }
        }
    }
}

/* calc_sat_mean_x / _z (LocalIntensityScale.hpp:365-494): running sum in a  */
/* ring of 2*off+1 lines; shared body over an outer axis of length n.        */
static inline void sat_mean_outer(float* base, int64_t n, int64_t s, int64_t y_num, int64_t off, std::vector<float>& tv) {
    const int64_t w = 2 * off + 1;
    for (int64_t k = 0; k < y_num; ++k) tv[k] = base[k];
    for (int64_t i = 1; i < w; ++i)
        for (int64_t k = 0; k < y_num; ++k) tv[i * y_num + k] = base[i * s + k] + tv[(i - 1) * y_num + k];
    for (int64_t i = 0; i < off + 1; ++i)
        for (int64_t k = 0; k < y_num; ++k) base[i * s + k] = tv[(i + off) * y_num + k] / (float)(i + off + 1);
    int64_t cur = off + 1, im = 0;
    for (int64_t i = off + 1; i < n - off; ++i) {
        im = (cur + off) % w;
        int64_t pm = (cur + off - 1) % w;
        for (int64_t k = 0; k < y_num; ++k) {
            float t = base[(i + off) * s + k] + tv[pm * y_num + k];
            base[i * s + k] = (t - tv[im * y_num + k]) / (float)w;
            tv[im * y_num + k] = t;
        }
        cur = (cur + 1) % w;
    }
    cur = (cur + off) % w;
    for (int64_t i = n - off; i < n; ++i) {
        for (int64_t k = 0; k < y_num; ++k) base[i * s + k] = (tv[im * y_num + k] - tv[cur * y_num + k]) / (float)(n - i + off);
        cur = (cur + 1) % w;
    }
}

static inline void sat_mean_x(Mesh<float>& m, int64_t off) {
    const int64_t z_num = m.z_num, x_num = m.x_num, y_num = m.y_num;
#pragma omp parallel
    {
        std::vector<float> tv(y_num * (2 * off + 1), 0.0f);
#pragma omp for schedule(static)
        for (int64_t j = 0; j < z_num; ++j) sat_mean_outer(m.v.data() + j * x_num * y_num, x_num, y_num, y_num, off, tv);
    }
}

static inline void sat_mean_z(Mesh<float>& m, int64_t off) {
    const int64_t z_num = m.z_num, x_num = m.x_num, y_num = m.y_num;
#pragma omp parallel
    {
        std::vector<float> tv(y_num * (2 * off + 1), 0.0f);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < x_num; ++i) sat_mean_outer(m.v.data() + i * y_num, z_num, x_num * y_num, y_num, off, tv);
    }
}

/* get_local_intensity_scale (LocalIntensityScale.hpp:26-150).               */
/* In: lis = downsampled smoothed intensities.  Out: lis = sigma,             */
/* intensity = copy of the input intensities (the Ip_th reference).           */
static inline void local_intensity_scale(Mesh<float>& lis, Mesh<float>& intensity, const aprh_params& par) {
    LisWindow w = lis_window(par, lis.y_num, lis.x_num, lis.z_num);
    const int nact = (int)w.active[0] + (int)w.active[1] + (int)w.active[2];
    if (par.constant_intensity_scale || nact == 0) {
        float min_val = 660000;
        double sum = 0;
        for (size_t i = 0; i < lis.size(); ++i) {
            float t = lis.v[i];
            sum += t;
            if (t < min_val) min_val = t;
        }
        float numel = (float)(lis.y_num * lis.x_num * lis.z_num);
        float sv = (float)(sum / (double)numel - (double)min_val);
        intensity = lis;  /* local_scale_temp2 untouched by the reference here; it holds the pre-LIS copy made by the caller */
        for (auto& v : lis.v) v = sv;
        return;
    }
    Mesh<float> copy = lis;
    if (par.reflect_bc_lis) {
        Mesh<float> a, b;
        padd(copy, a, std::max(w.win[0], w.win[3]), std::max(w.win[1], w.win[4]), std::max(w.win[2], w.win[5]));
        padd(copy, b, std::max(w.win[0], w.win[3]), std::max(w.win[1], w.win[4]), std::max(w.win[2], w.win[5]));
        if (w.active[0]) sat_mean_y(a, w.win[0]);
        if (w.active[1]) sat_mean_x(a, w.win[1]);
        if (w.active[2]) sat_mean_z(a, w.win[2]);
        for (size_t i = 0; i < a.size(); ++i) a.v[i] = std::fabs(a.v[i] - b.v[i]);
        if (w.active[0]) sat_mean_y(a, w.win[3]);
        if (w.active[1]) sat_mean_x(a, w.win[4]);
        if (w.active[2]) sat_mean_z(a, w.win[5]);
        for (size_t i = 0; i < a.size(); ++i) a.v[i] = a.v[i] * w.var_rescale;
        /* unpaddPixels (PixelData.hpp:1189-1210) */
        const int64_t py = (a.y_num - copy.y_num) / 2, px = (a.x_num - copy.x_num) / 2, pz = (a.z_num - copy.z_num) / 2;
        for (int64_t j = 0; j < copy.z_num; ++j)
            for (int64_t i = 0; i < copy.x_num; ++i)
                for (int64_t k = 0; k < copy.y_num; ++k) lis.at(k, i, j) = a.at(k + py, i + px, j + pz);
        intensity = copy;
    } else {
        Mesh<float> b = lis;
        Mesh<float>& a = lis;
        if (w.active[0]) sat_mean_y(a, w.win[0]);
        if (w.active[1]) sat_mean_x(a, w.win[1]);
        if (w.active[2]) sat_mean_z(a, w.win[2]);
        for (size_t i = 0; i < a.size(); ++i) a.v[i] = std::fabs(a.v[i] - b.v[i]);
        if (w.active[0]) sat_mean_y(a, w.win[3]);
        if (w.active[1]) sat_mean_x(a, w.win[4]);
        if (w.active[2]) sat_mean_z(a, w.win[5]);
        for (size_t i = 0; i < a.size(); ++i) a.v[i] = a.v[i] * w.var_rescale;
        intensity = b;
    }
}

/* ------------------------------------------------------------------------ */
/* GenInfo (GenInfo.hpp:33-65)                                               */
/* ------------------------------------------------------------------------ */
struct GenInfo {
    int l_min = 1, l_max = 1;
    int64_t org[3] = {0, 0, 0};
    std::vector<int64_t> y_num, x_num, z_num;
    void init(int64_t y, int64_t x, int64_t z) {
        org[0] = y; org[1] = x; org[2] = z;
        int64_t md = std::max(std::max(x, y), z);
        int lm = std::max(1, (int)std::ceil(std::log2((double)md)));
        l_min = 1;
        l_max = lm;
        y_num.assign(lm + 1, 0);
        x_num.assign(lm + 1, 0);
        z_num.assign(lm + 1, 0);
        for (int l = l_min; l <= l_max; ++l) {
            double cs = std::round(std::pow(2.0, (double)(l_max - l)));
            y_num[l] = (int64_t)std::ceil((double)y / cs);
            x_num[l] = (int64_t)std::ceil((double)x / cs);
            z_num[l] = (int64_t)std::ceil((double)z / cs);
        }
    }
    /* GenInfo::init_tree (GenInfo.hpp:68-99) */
    void init_tree(int64_t y, int64_t x, int64_t z) {
        org[0] = y; org[1] = x; org[2] = z;
        int64_t md = std::max(std::max(y, x), z);
        int lm = (int)std::ceil(std::log2((double)md));
        l_min = 0;
        l_max = std::max(lm - 1, 0);
        y_num.assign(l_max + 1, 0);
        x_num.assign(l_max + 1, 0);
        z_num.assign(l_max + 1, 0);
        for (int l = l_min; l <= l_max; ++l) {
            double cs = std::round(std::pow(2.0, (double)(l_max - l + 1)));
            y_num[l] = (int64_t)std::ceil((double)y / cs);
            x_num[l] = (int64_t)std::ceil((double)x / cs);
            z_num[l] = (int64_t)std::ceil((double)z / cs);
        }
    }
};

/* LinearAccess (LinearAccess.hpp:33-54) */
struct Linear {
    std::vector<uint64_t> level_xz_vec, xz_end_vec;
    std::vector<uint16_t> y_vec;
    uint64_t total = 0;
    void init_xz(const GenInfo& g) {
        uint64_t ct = 1;
        level_xz_vec.assign(g.l_max + 2, 0);
        level_xz_vec[0] = 1;
        for (int i = 0; i <= g.l_max; ++i) {
            ct += (uint64_t)(g.x_num[i] * g.z_num[i]);
            level_xz_vec[i + 1] = ct;
        }
        xz_end_vec.assign(ct, 0);
    }
    uint64_t row_begin(int level, int64_t z, int64_t x, const GenInfo& g) const {
        return xz_end_vec[level_xz_vec[level] + z * g.x_num[level] + x - 1];
    }
    uint64_t row_end(int level, int64_t z, int64_t x, const GenInfo& g) const {
        return xz_end_vec[level_xz_vec[level] + z * g.x_num[level] + x];
    }
};

/* ------------------------------------------------------------------------ */
/* Pulling scheme (PullingScheme.hpp:16-393)                                 */
/* ------------------------------------------------------------------------ */
enum : uint8_t { PS_EMPTY = 0, PS_SEED = 1, PS_BOUNDARY = 2, PS_FILLER = 3, PS_UPSEED = 4, PS_ASC = 8, PS_PROP = 15, PS_ASCN = 16 };

struct PullingScheme {
    int l_min = 1, l_max = 0;  /* l_max = GenInfo.l_max - 1 */
    std::vector<Mesh<uint8_t>> pct;

    void init(const GenInfo& g) {
        l_max = g.l_max - 1;
        l_min = g.l_min;
        pct.assign(std::max(l_max + 1, 0), Mesh<uint8_t>());
        for (int l = l_min; l <= l_max; ++l) {
            double cs = std::round(std::pow(2.0, (double)(l_max - l + 1)));
            pct[l].resize((int64_t)std::ceil((double)g.org[0] / cs), (int64_t)std::ceil((double)g.org[1] / cs),
                          (int64_t)std::ceil((double)g.org[2] / cs), PS_EMPTY);
        }
    }

    /* fill (PullingScheme.hpp:140-168) */
    void fill(int k, const Mesh<float>& in) {
        auto& m = pct[k].v;
        const float kf = (float)k;
        if (k == l_max) {
            for (size_t i = 0; i < in.size(); ++i) if (in.v[i] >= kf) m[i] = PS_SEED;
        } else if (k == l_min) {
            for (size_t i = 0; i < in.size(); ++i) if (in.v[i] <= kf) m[i] = PS_SEED;
        } else {
            for (size_t i = 0; i < in.size(); ++i) if (in.v[i] == kf) m[i] = PS_SEED;
        }
    }

    static void bounds(int64_t v, int64_t last, int64_t& lo, int64_t& hi) {
        lo = (v == 0) ? 0 : -1;
        hi = (v == last) ? 1 : 2;
    }

    /* set_ascendant_neighbours (PullingScheme.hpp:250-289) */
    void set_ascendant_neighbours(int level) {
        Mesh<uint8_t>& m = pct[level];
        for (int64_t j = 0; j < m.z_num; ++j) {
            int64_t zl, zh; bounds(j, m.z_num - 1, zl, zh);
            for (int64_t i = 0; i < m.x_num; ++i) {
                int64_t xl, xh; bounds(i, m.x_num - 1, xl, xh);
                for (int64_t k = 0; k < m.y_num; ++k) {
                    if (m.at(k, i, j) != PS_ASC) continue;
                    int64_t yl, yh; bounds(k, m.y_num - 1, yl, yh);
                    for (int64_t a = zl; a < zh; ++a)
                        for (int64_t b = xl; b < xh; ++b)
                            for (int64_t c = yl; c < yh; ++c) {
                                uint8_t& n = m.at(k + c, i + b, j + a);
                                if (n == PS_EMPTY) n = PS_ASCN;
                                if (n == PS_SEED) n = PS_PROP;
                            }
                }
            }
        }
    }

    /* set_filler (PullingScheme.hpp:291-343) */
    void set_filler(int level) {
        Mesh<uint8_t>& m = pct[level];
        Mesh<uint8_t>& c = pct[level + 1];
        for (int64_t j = 0; j < m.z_num; ++j) {
            int64_t cz = (j == m.z_num - 1 && (c.z_num % 2)) ? 1 : 2;
            for (int64_t i = 0; i < m.x_num; ++i) {
                int64_t cx = (i == m.x_num - 1 && (c.x_num % 2)) ? 1 : 2;
                for (int64_t k = 0; k < m.y_num; ++k) {
                    int64_t cy = (k == m.y_num - 1 && (c.y_num % 2)) ? 1 : 2;
                    uint8_t s = m.at(k, i, j);
                    if (s == PS_ASCN || s == PS_PROP) {
                        for (int64_t a = 2 * j; a < 2 * j + cz; ++a)
                            for (int64_t b = 2 * i; b < 2 * i + cx; ++b)
                                for (int64_t d = 2 * k; d < 2 * k + cy; ++d)
                                    if (c.at(d, b, a) == PS_EMPTY) c.at(d, b, a) = PS_FILLER;
                    }
                }
            }
        }
    }

    /* fill_neighbours + fill_parent (PullingScheme.hpp:345-393) */
    void fill_neighbours(int level) {
        Mesh<uint8_t>& m = pct[level];
        for (int64_t j = 0; j < m.z_num; ++j) {
            int64_t zl, zh; bounds(j, m.z_num - 1, zl, zh);
            for (int64_t i = 0; i < m.x_num; ++i) {
                int64_t xl, xh; bounds(i, m.x_num - 1, xl, xh);
                for (int64_t k = 0; k < m.y_num; ++k) {
                    uint8_t s = m.at(k, i, j);
                    if (s == PS_SEED || s == PS_PROP) {
                        int64_t yl, yh; bounds(k, m.y_num - 1, yl, yh);
                        for (int64_t a = zl; a < zh; ++a)
                            for (int64_t b = xl; b < xh; ++b)
                                for (int64_t c = yl; c < yh; ++c) {
                                    uint8_t& n = m.at(k + c, i + b, j + a);
                                    if (n == PS_EMPTY) n = PS_BOUNDARY;
                                }
                        fill_parent(j, i, k, m.x_num, m.y_num, level - 1);
                    } else if (s == PS_ASC) {
                        fill_parent(j, i, k, m.x_num, m.y_num, level - 1);
                    }
                }
            }
        }
    }

    void fill_parent(int64_t j, int64_t i, int64_t k, int64_t x_num, int64_t y_num, int nl) {
        if (nl >= l_min) {
            int64_t nx = (x_num + 1) / 2, ny = (y_num + 1) / 2;
            uint8_t& p = pct[nl].v[(size_t)((j / 2) * nx * ny + (i / 2) * ny + (k / 2))];
            if (p != PS_SEED) p = PS_ASC;
        }
    }

    /* pulling_scheme_main (PullingScheme.hpp:111-132) */
    void run() {
        for (int level = l_max; level >= l_min; --level) {
            if (level != l_max) {
                set_ascendant_neighbours(level);
                set_filler(level);
            }
            fill_neighbours(level);
        }
    }
};

/* computeLevels + compute_level_for_array + asmlog_2                        */
/* (LocalParticleCellSet.hpp:128-171). lis becomes the level map (as float). */
template <typename G>
static inline void compute_levels(const Mesh<G>& grad, Mesh<float>& lis, int max_level, float rel_error, float dx, float dy,
                                  float dz) {
    for (size_t i = 0; i < grad.size(); ++i) lis.v[i] = (float)grad.v[i] / lis.v[i];
    float min_dim = std::min(dy, std::min(dx, dz));
    float level_factor = (float)(std::pow(2.0, (double)max_level) * (double)min_dim);
    const float mult = level_factor / rel_error;
    for (size_t i = 0; i < lis.size(); ++i) {
        uint32_t u = f2u32(lis.v[i] * mult);
        lis.v[i] = (float)(u == 0 ? 0u : (uint32_t)(31 - __builtin_clz(u)));
    }
}

/* get_local_particle_cell_set (LocalParticleCellSet.hpp:58-91) */
static inline void local_particle_cell_set(PullingScheme& ps, Mesh<float>& levels) {
    ps.fill(ps.l_max, levels);
    Mesh<float> cur = levels, nxt;
    for (int l = ps.l_max - 1; l >= ps.l_min; --l) {
        downsample(cur, nxt, [](float a, float b) { return std::max(a, b); }, [](float a) { return a; });
        ps.fill(l, nxt);
        std::swap(cur, nxt);
    }
}

/* LinearAccess::initialize_linear_structure (LinearAccess.hpp:210-524) */
static inline void linear_structure(const GenInfo& g, Linear& L, std::vector<Mesh<uint8_t>>& pmap, bool nbhd_opt) {
    const uint8_t min_type = nbhd_opt ? 1 : 2;
    const uint8_t UPS = PS_UPSEED;
    L.init_xz(g);
    const int lmax = g.l_max, lmin = g.l_min;
    if (lmax <= 2) {
        const uint64_t ls = L.level_xz_vec[lmax];
        uint64_t counter = 0;
        for (int64_t z = 0; z < g.z_num[lmax]; ++z)
            for (int64_t x = 0; x < g.x_num[lmax]; ++x) {
                counter += (uint64_t)g.y_num[lmax];
                L.xz_end_vec[ls + z * g.x_num[lmax] + x] = counter;
            }
        L.y_vec.assign(counter, 0);
        counter = 0;
        for (int64_t z = 0; z < g.z_num[lmax]; ++z)
            for (int64_t x = 0; x < g.x_num[lmax]; ++x)
                for (int64_t y = 0; y < g.y_num[lmax]; ++y) L.y_vec[counter++] = (uint16_t)y;
        L.total = counter;
        return;
    }
    /* step 1: equivalence optimisation, levels processed upward in order */
    for (int level = lmin + 1; level < lmax; ++level) {
        const int64_t xL = g.x_num[level], zL = g.z_num[level], yL = g.y_num[level];
        const int64_t xU = g.x_num[level - 1], yU = g.y_num[level - 1];
        for (int64_t z = 0; z < zL; ++z)
            for (int64_t x = 0; x < xL; ++x) {
                const int64_t ods = (x / 2) * yU + (z / 2) * yU * xU, o = x * yL + z * yL * xL;
                for (int64_t y = 0; y < yU; ++y) {
                    uint8_t s = pmap[level - 1].v[ods + y];
                    if (s > 0 && s <= min_type) {
                        int64_t y2p = std::min(2 * y + 1, yL - 1);
                        pmap[level].v[o + 2 * y] = UPS;
                        pmap[level].v[o + y2p] = UPS;
                    }
                }
            }
    }
    /* step 2: row counts */
    for (int level = lmin; level < lmax - 1; ++level) {
        const int64_t xL = g.x_num[level], zL = g.z_num[level], yL = g.y_num[level];
        const uint64_t ls = L.level_xz_vec[level];
        for (int64_t z = 0; z < zL; ++z)
            for (int64_t x = 0; x < xL; ++x) {
                const int64_t opc = z * xL + x, opm = yL * opc;
                uint64_t c = 0;
                for (int64_t y = 0; y < yL; ++y) {
                    uint8_t s = pmap[level].v[opm + y];
                    if (s > min_type && s <= UPS) c++;
                }
                L.xz_end_vec[ls + opc] = c;
            }
    }
    const int lm1 = lmax - 1;
    const int64_t xL = g.x_num[lm1], zL = g.z_num[lm1], yL = g.y_num[lm1], yLm = g.y_num[lm1 + 1];
    const uint64_t ls1 = L.level_xz_vec[lm1];
    std::vector<uint64_t> tmax((size_t)(zL * xL), 0);
    for (int64_t z = 0; z < zL; ++z)
        for (int64_t x = 0; x < xL; ++x) {
            const int64_t opc = z * xL + x, opm = yL * opc;
            uint64_t c = 0, cl = 0;
            for (int64_t y = 0; y < yL; ++y) {
                uint8_t s = pmap[lm1].v[opm + y];
                if (s > min_type && s <= UPS) c++;
                else if (s > 0 && s <= min_type) {
                    cl++;
                    if (2 * y < yLm - 1) cl++;
                }
            }
            L.xz_end_vec[ls1 + opc] = c;
            tmax[opc] = cl;
        }
    const int64_t xLm = g.x_num[lmax], zLm = g.z_num[lmax];
    const uint64_t lsm = L.level_xz_vec[lmax];
    for (int64_t z = 0; z < zLm; ++z)
        for (int64_t x = 0; x < xLm; ++x) L.xz_end_vec[lsm + z * xLm + x] = tmax[(z / 2) * xL + x / 2];
    std::partial_sum(L.xz_end_vec.begin(), L.xz_end_vec.end(), L.xz_end_vec.begin());
    L.total = L.xz_end_vec.back();
    L.y_vec.assign(L.total, 0);
    for (int level = lmin; level < lmax - 1; ++level) {
        const int64_t xl = g.x_num[level], zl = g.z_num[level], yl = g.y_num[level];
        const uint64_t ls = L.level_xz_vec[level];
        for (int64_t z = 0; z < zl; ++z)
            for (int64_t x = 0; x < xl; ++x) {
                const int64_t opc = z * xl + x, opm = yl * opc;
                uint64_t c = 0, oy = L.xz_end_vec[ls + opc - 1];
                for (int64_t y = 0; y < yl; ++y) {
                    uint8_t s = pmap[level].v[opm + y];
                    if (s > min_type && s <= UPS) L.y_vec[oy + c++] = (uint16_t)y;
                }
            }
    }
    for (int64_t z = 0; z < zL; ++z)
        for (int64_t x = 0; x < xL; ++x) {
            const int64_t opc = z * xL + x, opcm = (z * 2) * xLm + x * 2, opm = yL * opc;
            uint64_t c = 0, cl = 0;
            uint64_t oy = L.xz_end_vec[ls1 + opc - 1], oym = L.xz_end_vec[lsm + opcm - 1];
            for (int64_t y = 0; y < yL; ++y) {
                uint8_t s = pmap[lm1].v[opm + y];
                if (s > min_type && s <= UPS) L.y_vec[oy + c++] = (uint16_t)y;
                else if (s > 0 && s <= min_type) {
                    L.y_vec[oym + cl++] = (uint16_t)(2 * y);
                    if (2 * y < yLm - 1) L.y_vec[oym + cl++] = (uint16_t)(2 * y + 1);
                }
            }
        }
    for (int64_t z = 0; z < zLm; ++z)
        for (int64_t x = 0; x < xLm; ++x) {
            if ((z % 2) != 0 || (x % 2) != 0) {
                const int64_t om = z * xLm + x, omf = (z / 2) * 2 * xLm + (x / 2) * 2;
                uint64_t bf = L.xz_end_vec[lsm + omf - 1], ef = L.xz_end_vec[lsm + omf], b = L.xz_end_vec[lsm + om - 1];
                std::copy(L.y_vec.begin() + bf, L.y_vec.begin() + ef, L.y_vec.begin() + b);
            }
        }
}

/* ------------------------------------------------------------------------ */
/* get_apr (APRConverter.hpp:534-582) for ImageType = uint16_t               */
/* ------------------------------------------------------------------------ */
struct AprResult {
    GenInfo info;
    Linear lin;
    float bspline_offset = 0;
    /* intermediates (for stage-level parity) */
    Mesh<uint16_t> smoothed;   /* after the three IIR passes            */
    Mesh<uint16_t> grad;       /* after thresholds (input to levels)     */
    Mesh<uint16_t> grad_raw;   /* gradient_step.tif equivalent           */
    bool is_float = false;     /* APRConverter<float>: the three above as float */
    Mesh<float> smoothed_f, grad_f, grad_raw_f;
    Mesh<float> lis_raw;       /* local_intensity_scale_step.tif equiv.  */
    Mesh<float> lis;           /* after sigma_th rescale                 */
    Mesh<float> levels;        /* computeLevels output                   */
    std::vector<Mesh<uint8_t>> pct;  /* final particle-cell tree        */
};

/* compute_bspline_offset (AutoParameters.hpp:165-196) for uint16 */
static inline float bspline_offset_u16(const uint16_t* img, size_t n, float lambda) {
    if (lambda <= 0) return 0;
    uint16_t mn = 65535, mx = 0;
    for (size_t i = 0; i < n; ++i) {
        mn = std::min(mn, img[i]);
        mx = std::max(mx, img[i]);
    }
    float off = (float)((65535 - ((int)mx - (int)mn)) / 2 - (int)mn);
    return std::max(std::min(off, 100.f), -100.f);
}

/* The stages after computeL: applyParameters, solveForAPR,                  */
/* generateDatastructures (APRConverter.hpp:249-363).                        */
template <typename G>
static inline void keep_grad(AprResult& r, const Mesh<G>& g);
template <>
inline void keep_grad<uint16_t>(AprResult& r, const Mesh<uint16_t>& g) { r.grad = g; }
template <>
inline void keep_grad<float>(AprResult& r, const Mesh<float>& g) { r.grad_f = g; }

template <typename G>
static inline void apr_from_grad_lis(AprResult& r, Mesh<G>& grad, Mesh<float>& lis, Mesh<float>& intensity,
                                     const aprh_params& par, bool keep) {
    /* applyParameters (APRConverter.hpp:249-297) */
    const float ipth = par.Ip_th + r.bspline_offset;
    for (size_t i = 0; i < intensity.size(); ++i) if (intensity.v[i] <= ipth) grad.v[i] = 0;
    for (size_t i = 0; i < grad.size(); ++i) {
        float s = lis.v[i];
        if (s < par.sigma_th) {
            s = (s < par.sigma_th_max) ? 60000.0f : par.sigma_th;
            lis.v[i] = s;
        }
    }
    for (size_t i = 0; i < grad.size(); ++i) if ((float)grad.v[i] < par.grad_th) grad.v[i] = 0;
    if (keep) { keep_grad(r, grad); r.lis = lis; }
    /* solveForAPR (APRConverter.hpp:300-336) */
    compute_levels(grad, lis, r.info.l_max, par.rel_error, par.dx, par.dy, par.dz);
    if (keep) r.levels = lis;
    PullingScheme ps;
    ps.init(r.info);
    local_particle_cell_set(ps, lis);
    ps.run();
    /* generateDatastructures (APRConverter.hpp:338-363) */
    linear_structure(r.info, r.lin, ps.pct, par.neighborhood_optimization != 0);
    if (keep) r.pct = ps.pct;
}

static inline void get_apr_u16(AprResult& r, const Mesh<uint16_t>& input, const aprh_params& par, bool keep) {
    r.info.init(input.y_num, input.x_num, input.z_num);
    const int64_t zd = (int64_t)std::ceil(1.0 * input.z_num / 2.0), xd = (int64_t)std::ceil(1.0 * input.x_num / 2.0),
                  yd = (int64_t)std::ceil(1.0 * input.y_num / 2.0);
    Mesh<uint16_t> grad(yd, xd, zd, 0);
    /* computeL (APRConverter.hpp:194-247) */
    Mesh<uint16_t> img(input.y_num, input.x_num, input.z_num);
    r.bspline_offset = bspline_offset_u16(input.v.data(), input.size(), par.lambda);
    const float off = r.bspline_offset;
    for (size_t i = 0; i < input.size(); ++i) img.v[i] = f2u16((float)input.v[i] + off);
    /* ComputeGradient::get_gradient (ComputeGradient.hpp:70-119) */
    if (par.lambda > 0) smooth_bspline_3d(img, par.lambda);
    if (keep) r.smoothed = img;
    grad_mag_ds(img, grad, par.dx, par.dy, par.dz);
    Mesh<float> lis;
    downsample(img, lis, [](float a, float b) { return a + b; }, [](float a) { return (float)((double)a / 8.0); });
    if (par.lambda > 0) {
        if (img.y_num > 2) inv_bspline_y(lis);
        if (img.x_num > 2) inv_bspline_x(lis);
        if (img.z_num > 2) inv_bspline_z(lis);
    }
    if (keep) r.grad_raw = grad;
    Mesh<float> intensity;
    local_intensity_scale(lis, intensity, par);
    if (keep) r.lis_raw = lis;
    apr_from_grad_lis(r, grad, lis, intensity, par, keep);
}

/* ------------------------------------------------------------------------ */
/* Blocked conversion: APRConverterBatch::get_apr_batch_method_from_file     */
/* (APRConverterBatch.hpp:111-175) on an image held in memory.  z blocks of  */
/* z_block_size planes (the last takes the remainder) with ghost slabs of up */
/* to ghost_z planes; each patch runs the pixel pipeline on its own (its own */
/* B-spline offset, get_apr_method_patch :186-276) with the global l_max in  */
/* computeLevels, and fills the global particle-cell tree over its non-ghost */
/* planes (fill_global_particle_cell_tree :287-322, PullingScheme::fill_patch */
/* PullingScheme.hpp:175-244); then one pulling scheme and linear structure. */
/* ------------------------------------------------------------------------ */
struct ZBlock {
    int64_t z0, zf, gl, gr;   /* block [z0, zf), ghosts below / above */
};

static inline std::vector<ZBlock> z_blocks(int64_t z_num, int64_t z_block_size, int64_t ghost_z) {
    const int64_t nb = std::max<int64_t>(z_num / z_block_size, 1);
    std::vector<ZBlock> out;
    for (int64_t b = 0; b < nb; ++b) {
        ZBlock k;
        k.z0 = b * z_block_size;
        k.zf = (b == nb - 1) ? z_num : (b + 1) * z_block_size;
        k.gl = std::min(k.z0, ghost_z);
        k.gr = std::min(z_num - k.zf, ghost_z);
        out.push_back(k);
    }
    return out;
}

/* fill_patch: SEED where the patch's level map at pct level k satisfies the level's test, over
   the patch planes [gl / f, z_num - gr / f) moved by z0 / f - gl / f (f = 2^(l_max_pct + 1 - k)) */
static inline void fill_patch(PullingScheme& ps, int k, const Mesh<float>& in, const ZBlock& b) {
    const int64_t f = (int64_t)1 << (ps.l_max + 1 - k);
    const int64_t zl = b.gl / f, zr = b.gr / f;
    const int64_t oz = (b.z0 - b.gl + b.gl) / f - zl;
    Mesh<uint8_t>& m = ps.pct[k];
    const float kf = (float)k;
    for (int64_t z = zl; z < in.z_num - zr; ++z) {
        if (z + oz >= m.z_num) continue;
        for (int64_t x = 0; x < in.x_num; ++x)
            for (int64_t y = 0; y < in.y_num; ++y) {
                const float v = in.at(y, x, z);
                const bool seed = (k == ps.l_max) ? (v >= kf) : (k == ps.l_min) ? (v <= kf) : (v == kf);
                if (seed) m.at(y, x, z + oz) = PS_SEED;
            }
    }
}

static inline void get_apr_by_block_u16(AprResult& r, const Mesh<uint16_t>& input, const aprh_params& par, int64_t z_block_size,
                                        int64_t ghost_z) {
    r.info.init(input.y_num, input.x_num, input.z_num);
    PullingScheme ps;
    ps.init(r.info);
    const int64_t plane = input.y_num * input.x_num;
    for (const ZBlock& b : z_blocks(input.z_num, z_block_size, ghost_z)) {
        const int64_t zb = b.z0 - b.gl, ze = b.zf + b.gr;
        Mesh<uint16_t> patch(input.y_num, input.x_num, ze - zb);
        std::copy(input.v.begin() + zb * plane, input.v.begin() + ze * plane, patch.v.begin());
        const int64_t zd = (int64_t)std::ceil(1.0 * patch.z_num / 2.0), xd = (int64_t)std::ceil(1.0 * patch.x_num / 2.0),
                      yd = (int64_t)std::ceil(1.0 * patch.y_num / 2.0);
        /* get_apr_method_patch: the pixel pipeline of get_apr_u16 on the patch */
        Mesh<uint16_t> grad(yd, xd, zd, 0);
        Mesh<uint16_t> img(patch.y_num, patch.x_num, patch.z_num);
        const float off = bspline_offset_u16(patch.v.data(), patch.size(), par.lambda);
        for (size_t i = 0; i < patch.size(); ++i) img.v[i] = f2u16((float)patch.v[i] + off);
        if (par.lambda > 0) smooth_bspline_3d(img, par.lambda);
        grad_mag_ds(img, grad, par.dx, par.dy, par.dz);
        Mesh<float> lis;
        downsample(img, lis, [](float a, float c) { return a + c; }, [](float a) { return (float)((double)a / 8.0); });
        if (par.lambda > 0) {
            if (img.y_num > 2) inv_bspline_y(lis);
            if (img.x_num > 2) inv_bspline_x(lis);
            if (img.z_num > 2) inv_bspline_z(lis);
        }
        Mesh<float> intensity;
        local_intensity_scale(lis, intensity, par);
        /* applyParameters (APRConverterBatch.hpp:396-433) */
        const float ipth = par.Ip_th + off;
        for (size_t i = 0; i < intensity.size(); ++i) if (intensity.v[i] <= ipth) grad.v[i] = 0;
        for (size_t i = 0; i < grad.size(); ++i) {
            float s = lis.v[i];
            if (s < par.sigma_th) {
                s = (s < par.sigma_th_max) ? 60000.0f : par.sigma_th;
                lis.v[i] = s;
            }
        }
        for (size_t i = 0; i < grad.size(); ++i) if ((float)grad.v[i] < par.grad_th) grad.v[i] = 0;
        compute_levels(grad, lis, r.info.l_max, par.rel_error, par.dx, par.dy, par.dz);
        /* fill_global_particle_cell_tree */
        fill_patch(ps, ps.l_max, lis, b);
        Mesh<float> cur = lis, nxt;
        for (int l = ps.l_max - 1; l >= ps.l_min; --l) {
            downsample(cur, nxt, [](float a, float c) { return std::max(a, c); }, [](float a) { return a; });
            fill_patch(ps, l, nxt, b);
            std::swap(cur, nxt);
        }
    }
    ps.run();
    linear_structure(r.info, r.lin, ps.pct, par.neighborhood_optimization != 0);
}

/* ParticleData::sample_parts_from_img_blocked (ParticleData.hpp:292-362): per z block the mean
   pyramid of the patch (ghosts included) is sampled at the rows of the block's non-ghost planes
   of each level (f = 2^(l_max - level), planes [gl / f, z_num - gr / f) moved by z0 / f - gl / f);
   particles of rows no block reaches stay 0 */
template <typename ImageType, typename PartType>
void sample_image_blocked(const GenInfo& g, const Linear& L, const Mesh<ImageType>& img, int64_t z_block_size, int64_t ghost_z,
                          std::vector<PartType>& parts) {
    parts.assign(L.total, 0);
    const int64_t plane = img.y_num * img.x_num;
    for (const ZBlock& b : z_blocks(img.z_num, z_block_size, ghost_z)) {
        const int64_t zb = b.z0 - b.gl, ze = b.zf + b.gr;
        std::vector<Mesh<ImageType>> pyr(g.l_max + 1);
        pyr[g.l_max].resize(img.y_num, img.x_num, ze - zb);
        std::copy(img.v.begin() + zb * plane, img.v.begin() + ze * plane, pyr[g.l_max].v.begin());
        for (int level = g.l_max; level > g.l_min; --level)
            downsample(pyr[level], pyr[level - 1], [](float a, float c) { return a + c; },
                       [](float a) { return Store<ImageType>::cvt(a / 8.0f); });
        for (int level = g.l_min; level <= g.l_max; ++level) {
            const int64_t f = (int64_t)1 << (g.l_max - level);
            const int64_t zl = b.gl / f, zr = b.gr / f;
            const int64_t oz = (b.z0 - b.gl + b.gl) / f - zl;
            const Mesh<ImageType>& P = pyr[level];
            for (int64_t z = zl; z < P.z_num - zr; ++z) {
                if (z + oz >= g.z_num[level]) continue;
                for (int64_t x = 0; x < P.x_num; ++x)
                    for (uint64_t i = L.row_begin(level, z + oz, x, g); i < L.row_end(level, z + oz, x, g); ++i)
                        if ((int64_t)L.y_vec[i] < P.y_num) parts[i] = (PartType)P.at(L.y_vec[i], x, z);
            }
        }
    }
}

/* get_apr for ImageType = float (APRConverter<float>): no B-spline offset   */
/* (APRConverter.hpp:218-220), float smoothing, float gradient (grad_temp is */
/* PixelData<float>), then the same stages.                                  */
static inline void get_apr_f32(AprResult& r, const Mesh<float>& input, const aprh_params& par, bool keep) {
    r.info.init(input.y_num, input.x_num, input.z_num);
    r.is_float = true;
    const int64_t zd = (int64_t)std::ceil(1.0 * input.z_num / 2.0), xd = (int64_t)std::ceil(1.0 * input.x_num / 2.0),
                  yd = (int64_t)std::ceil(1.0 * input.y_num / 2.0);
    Mesh<float> grad(yd, xd, zd, 0.0f);
    Mesh<float> img = input;
    r.bspline_offset = 0;
    if (par.lambda > 0) smooth_bspline_3d(img, par.lambda);
    if (keep) r.smoothed_f = img;
    grad_mag_ds(img, grad, par.dx, par.dy, par.dz);
    Mesh<float> lis;
    downsample(img, lis, [](float a, float b) { return a + b; }, [](float a) { return (float)((double)a / 8.0); });
    if (par.lambda > 0) {
        if (img.y_num > 2) inv_bspline_y(lis);
        if (img.x_num > 2) inv_bspline_x(lis);
        if (img.z_num > 2) inv_bspline_z(lis);
    }
    if (keep) r.grad_raw_f = grad;
    Mesh<float> intensity;
    local_intensity_scale(lis, intensity, par);
    if (keep) r.lis_raw = lis;
    apr_from_grad_lis(r, grad, lis, intensity, par, keep);
}

/* ParticleData::sample_image (ParticleData.hpp:122-127, :221-261) with the  */
/* mean pyramid of downsamplePyramid (PixelData.hpp:1036-1058).              */
template <typename ImageType, typename PartType>
void sample_image(const GenInfo& g, const Linear& L, const Mesh<ImageType>& img, std::vector<PartType>& parts) {
    std::vector<Mesh<ImageType>> pyr(g.l_max + 1);
    pyr[g.l_max] = img;
    for (int level = g.l_max; level > g.l_min; --level)
        downsample(pyr[level], pyr[level - 1], [](float a, float b) { return a + b; },
                   [](float a) { return Store<ImageType>::cvt(a / 8.0f); });
    parts.assign(L.total, 0);
    for (int level = g.l_min; level <= g.l_max; ++level) {
#pragma omp parallel for schedule(dynamic)
        for (int64_t z = 0; z < g.z_num[level]; ++z)
            for (int64_t x = 0; x < g.x_num[level]; ++x)
                for (uint64_t i = L.row_begin(level, z, x, g); i < L.row_end(level, z, x, g); ++i)
                    parts[i] = (PartType)pyr[level].at(L.y_vec[i], x, z);
    }
}


/* ======================================================================== */
/* Particle operations on an APR: tree, tree means, stencils, convolution,   */
/* Richardson-Lucy.                                                          */
/* ======================================================================== */

struct Tree {
    GenInfo info;   /* tree geometry (GenInfo::init_tree)            */
    Linear lin;     /* tree LinearAccess (interior cells)             */
};

/* APR::initialize_tree_dense (APR.hpp:633-689) +                            */
/* LinearAccess::initialize_tree_access_dense (LinearAccess.hpp:67-122)      */
static inline void build_tree(const GenInfo& g, const Linear& L, Tree& t) {
    t.info.init_tree(g.org[0], g.org[1], g.org[2]);
    const GenInfo& ti = t.info;
    std::vector<Mesh<uint8_t>> m(ti.l_max + 1);
    for (int l = ti.l_min; l <= ti.l_max; ++l) m[l].resize(ti.y_num[l], ti.x_num[l], ti.z_num[l], 0);
    /* parents of APR particles */
    for (int level = g.l_max; level >= g.l_min; --level) {
        if (level - 1 < ti.l_min || level - 1 > ti.l_max) continue;
        for (int64_t z = 0; z < g.z_num[level]; ++z)
            for (int64_t x = 0; x < g.x_num[level]; ++x)
                for (uint64_t i = L.row_begin(level, z, x, g); i < L.row_end(level, z, x, g); ++i)
                    m[level - 1].at(L.y_vec[i] / 2, x / 2, z / 2) = 1;
    }
    /* the rest of the tree, level by level ("suffices to check one child") */
    for (int level = ti.l_max - 1; level >= ti.l_min; --level)
        for (int64_t z = 0; z < ti.z_num[level]; ++z)
            for (int64_t x = 0; x < ti.x_num[level]; ++x)
                for (int64_t y = 0; y < ti.y_num[level]; ++y)
                    m[level].at(y, x, z) = m[level].at(y, x, z) | m[level + 1].at(2 * y, 2 * x, 2 * z);
    /* linear access of the tree */
    t.lin.init_xz(ti);
    for (int level = ti.l_max; level >= ti.l_min; --level)
        for (int64_t z = 0; z < ti.z_num[level]; ++z)
            for (int64_t x = 0; x < ti.x_num[level]; ++x) {
                uint64_t c = 0;
                for (int64_t y = 0; y < ti.y_num[level]; ++y) c += m[level].at(y, x, z);
                t.lin.xz_end_vec[t.lin.level_xz_vec[level] + z * ti.x_num[level] + x] = c;
            }
    std::partial_sum(t.lin.xz_end_vec.begin(), t.lin.xz_end_vec.end(), t.lin.xz_end_vec.begin());
    t.lin.total = t.lin.xz_end_vec.back();
    t.lin.y_vec.assign(t.lin.total, 0);
    for (int level = ti.l_max; level >= ti.l_min; --level)
        for (int64_t z = 0; z < ti.z_num[level]; ++z)
            for (int64_t x = 0; x < ti.x_num[level]; ++x) {
                uint64_t pi = t.lin.row_begin(level, z, x, ti);
                for (int64_t y = 0; y < ti.y_num[level]; ++y)
                    if (m[level].at(y, x, z)) t.lin.y_vec[pi++] = (uint16_t)y;
            }
}

/* APRTreeNumerics::fill_tree_mean (APRTreeNumerics.hpp:63-202): tree value =
   sum over children of scale*value/8, accumulated in particle order: all
   particle children first (levels l_max..l_min), then tree children
   (tree levels from the top down).                                          */
template <typename PartT>
void fill_tree_mean(const GenInfo& g, const Linear& L, const Tree& t, const PartT* parts, std::vector<float>& tree) {
    const GenInfo& ti = t.info;
    tree.assign(t.lin.total, 0.0f);
    auto accumulate = [&](int level, const GenInfo& gc, const Linear& Lc, auto value_of) {
        /* children at `level` (geometry gc / Lc), parents at tree level `level - 1` */
        const int pl = level - 1;
        const int64_t pxn = ti.x_num[pl], pzn = ti.z_num[pl], pyn = ti.y_num[pl];
#pragma omp parallel for schedule(dynamic)
        for (int64_t zd = 0; zd < pzn; ++zd) {
            for (int64_t z = 2 * zd; z <= std::min(2 * zd + 1, gc.z_num[level] - 1); ++z) {
                for (int64_t xd = 0; xd < pxn; ++xd) {
                    for (int64_t x = 2 * xd; x <= std::min(2 * xd + 1, gc.x_num[level] - 1); ++x) {
                        uint64_t pb = t.lin.row_begin(pl, z / 2, x / 2, ti), pe = t.lin.row_end(pl, z / 2, x / 2, ti);
                        float sxz = (float)((((2 * pxn != gc.x_num[level]) && ((x / 2) == (pxn - 1))) +
                                             ((2 * pzn != gc.z_num[level]) && ((z / 2) == (pzn - 1)))) * 2);
                        if (sxz == 0) sxz = 1;
                        float syxz = sxz;
                        if (2 * pyn != gc.y_num[level]) syxz = sxz * 2;
                        uint64_t pi = pb;
                        for (uint64_t i = Lc.row_begin(level, z, x, gc); i < Lc.row_end(level, z, x, gc); ++i) {
                            const int64_t y = Lc.y_vec[i];
                            while (pi < pe && t.lin.y_vec[pi] != y / 2) ++pi;
                            const float v = value_of(i);
                            if ((int64_t)t.lin.y_vec[pi] == pyn - 1) tree[pi] = syxz * v / 8.0f + tree[pi];
                            else tree[pi] = sxz * v / 8.0f + tree[pi];
                        }
                    }
                }
            }
        }
    };
    for (int level = g.l_max; level >= g.l_min; --level) {
        if (level - 1 < ti.l_min) continue;
        accumulate(level, g, L, [&](uint64_t i) { return (float)parts[i]; });
    }
    for (int level = ti.l_max; level > ti.l_min; --level)
        accumulate(level, ti, t.lin, [&](uint64_t i) { return tree[i]; });
}

/* APRStencil::downsample_stencil / get_downsampled_stencils (APRStencil.hpp:12-95) */
static inline void downsample_stencil(const Mesh<float>& in, Mesh<float>& out, int level_delta, bool normalize) {
    const int z_num = (int)in.z_num, x_num = (int)in.x_num, y_num = (int)in.y_num;
    const int ndim = (z_num > 1) + (x_num > 1) + (y_num > 1);
    const int step = (int)std::pow(2.0f, (float)level_delta);
    const int factor = (int)std::pow((float)step, (float)ndim);
    auto dsz = [&](int n) { return n > 1 ? std::max(2 * (((n - 1) / 2 + step - 1) / step) + 1, 3) : 1; };
    const int zd = dsz(z_num), xd = dsz(x_num), yd = dsz(y_num);
    out.resize(yd, xd, zd, 0.0f);
    const int zo = (zd / 2) * step - z_num / 2, xo = (xd / 2) * step - x_num / 2, yo = (yd / 2) * step - y_num / 2;
    for (int iz = 0; iz < z_num; ++iz)
        for (int ix = 0; ix < x_num; ++ix)
            for (int iy = 0; iy < y_num; ++iy)
                for (int dz = 0; dz < zd; ++dz) {
                    const float ovz = (float)std::max(std::min(dz * step + step, zo + iz + step) - std::max(dz * step, zo + iz), 0);
                    for (int dx = 0; dx < xd; ++dx) {
                        const float ovx = (float)std::max(std::min(dx * step + step, xo + ix + step) - std::max(dx * step, xo + ix), 0);
                        for (int dy = 0; dy < yd; ++dy) {
                            const float ovy = (float)std::max(std::min(dy * step + step, yo + iy + step) - std::max(dy * step, yo + iy), 0);
                            const float ov = ovx * ovy * ovz;
                            out.at(dy, dx, dz) += ov * in.at(iy, ix, iz);
                        }
                    }
                }
    float sum = 0;
    for (auto& v : out.v) {
        v /= (float)factor;
        sum += v;
    }
    if (normalize) {
        const float nf = 1.0f / sum;
        for (auto& v : out.v) v *= nf;
    }
}

/* PixelData::normalize (PixelData.hpp:771-781): divide by the float sum */
static inline void normalize_mesh(Mesh<float>& m) {
    float sum = 0;
    for (float v : m.v) sum += v;
    for (auto& v : m.v) v /= sum;
}

static inline void downsampled_stencils(const Mesh<float>& in, std::vector<Mesh<float>>& out, int nlevels, bool normalize) {
    out.assign(std::max(nlevels, 1), Mesh<float>());
    out[0] = in;
    if (normalize) normalize_mesh(out[0]);
    for (int d = 1; d < nlevels; ++d) downsample_stencil(in, out[d], d, normalize);
}

/* APRFilter::convolve_pencil (APRFilter.hpp:1094-1216) with its dense ring   */
/* buffer (Sz x Sx x y) and boundary helpers (:384-504, :705-849).           */
struct RingBuf {
    int64_t y_num = 0, x_num = 0, z_num = 0;
    std::vector<float> v;
    float& at(int64_t y, int64_t x, int64_t z) { return v[(size_t)(z * x_num * y_num + x * y_num + y)]; }
};

template <typename PartT>
void convolve_pencil(const GenInfo& g, const Linear& L, const Tree& t, const std::vector<Mesh<float>>& stencils, const PartT* in,
                     const float* tree_data, float* out, bool reflect) {
    const int nthreads =
#ifdef _OPENMP
        omp_get_max_threads();
#else
        1;
#endif
    int64_t ynm = (g.org[0] > 1) ? g.y_num[g.l_max] + stencils[0].y_num - 1 : 1;
    std::vector<RingBuf> bufs(nthreads);
    for (auto& b : bufs) b.v.assign((size_t)(ynm * stencils[0].x_num * stencils[0].z_num), 0.0f);
    const GenInfo& ti = t.info;
    for (int level = g.l_max; level >= g.l_min; --level) {
        const int sn = std::min((int)stencils.size() - 1, g.l_max - level);
        const Mesh<float>& st = stencils[sn];
        const int sy = (int)st.y_num, sx = (int)st.x_num, sz = (int)st.z_num;
        const int hy = (sy - 1) / 2, hx = (sx - 1) / 2, hz = (sz - 1) / 2;
        const int mr = std::max(hy, std::max(hx, hz));
        const int npl_max = (int)std::ceil(std::log2(mr + 2) - 1);
        const int npl = std::min(npl_max, level - g.l_min);
        const int64_t z_num = g.z_num[level], x_num = g.x_num[level];
        ynm = (g.org[0] > 1) ? g.y_num[level] + sy - 1 : 1;
        for (auto& b : bufs) { b.y_num = ynm; b.x_num = sx; b.z_num = sz; }

        auto bc_y = [&](RingBuf& b, int64_t z, int64_t x) {
            const int64_t base = (z % sz) * b.x_num * b.y_num + (x % sx) * b.y_num;
            if (reflect) {
                for (int y = 0; y < hy; ++y) b.v[base + hy - 1 - y] = b.v[base + hy + 1 + y];
                for (int y = 0; y < hy; ++y) {
                    const int64_t r = b.y_num - 1 - hy;
                    b.v[base + r + 1 + y] = b.v[base + r - 1 - y];
                }
            } else {
                for (int y = 0; y < hy; ++y) b.v[base + y] = 0;
                for (int y = 0; y < hy; ++y) b.v[base + b.y_num - hy + y] = 0;
            }
        };
        auto bc_z = [&](RingBuf& b, int64_t z, bool low) {
            const int64_t slice = b.x_num * b.y_num;
            const int64_t oo = (z % sz) * slice;
            if (reflect) {
                int64_t zin;
                if (low) zin = (uint64_t)(hz + (hz - z)) % (uint64_t)sz;
                else {
                    const uint64_t r = (uint64_t)(z_num - 1 + hz);
                    zin = (int64_t)((r - ((uint64_t)z - r)) % (uint64_t)sz);
                }
                std::copy(b.v.begin() + zin * slice, b.v.begin() + zin * slice + slice, b.v.begin() + oo);
            } else {
                std::fill(b.v.begin() + oo, b.v.begin() + oo + slice, 0.0f);
            }
        };
        auto bc_x = [&](RingBuf& b, int64_t x, bool low) {
            for (int zz = 0; zz < sz; ++zz) {
                const int64_t oo = zz * b.x_num * b.y_num + (x % sx) * b.y_num;
                if (reflect) {
                    int64_t xin;
                    if (low) xin = (uint64_t)(hx + (hx - x)) % (uint64_t)sx;
                    else {
                        const uint64_t r = (uint64_t)(x_num - 1 + hx);
                        xin = (int64_t)((r - ((uint64_t)x - r)) % (uint64_t)sx);
                    }
                    const int64_t io = zz * b.x_num * b.y_num + xin * b.y_num;
                    std::copy(b.v.begin() + io, b.v.begin() + io + b.y_num, b.v.begin() + oo);
                } else {
                    std::fill(b.v.begin() + oo, b.v.begin() + oo + b.y_num, 0.0f);
                }
            }
        };
        auto update = [&](RingBuf& b, int64_t z, int64_t x) {
            const int64_t mo = ((z + hz) % sz) * b.x_num * b.y_num + ((x + hx) % sx) * b.y_num + hy;
            for (uint64_t i = L.row_begin(level, z, x, g); i < L.row_end(level, z, x, g); ++i) b.v[mo + L.y_vec[i]] = (float)in[i];
            if (level > g.l_min) {
                for (int d = 1; d <= npl; ++d) {
                    const int step = (int)std::pow(2, d);
                    for (uint64_t i = L.row_begin(level - d, z / step, x / step, g); i < L.row_end(level - d, z / step, x / step, g); ++i) {
                        const int64_t y0 = step * L.y_vec[i];
                        const int64_t y1 = std::min<int64_t>(y0 + step, g.y_num[level]);
                        for (int64_t y = y0; y < y1; ++y) b.v[mo + y] = (float)in[i];
                    }
                }
            }
            if (level < g.l_max) {
                for (uint64_t i = t.lin.row_begin(level, z, x, ti); i < t.lin.row_end(level, z, x, ti); ++i)
                    b.v[mo + t.lin.y_vec[i]] = tree_data[i];
            }
            bc_y(b, z + hz, x + hx);
        };
        auto run = [&](RingBuf& b, int64_t z, int64_t x) {
            const int64_t xy = b.x_num * b.y_num;
            for (uint64_t i = L.row_begin(level, z, x, g); i < L.row_end(level, z, x, g); ++i) {
                float val = 0;
                const int64_t y = L.y_vec[i];
                size_t c = 0;
                for (int iz = 0; iz < sz; ++iz) {
                    const int64_t bo = ((z + iz) % sz) * xy + y;
                    for (int ix = 0; ix < sx; ++ix) {
                        const int64_t o = bo + ((x + ix) % sx) * b.y_num;
                        for (int iy = 0; iy < sy; ++iy) val += b.v[o + iy] * st.v[c++];
                    }
                }
                out[i] = val;
            }
        };
#pragma omp parallel for schedule(dynamic)
        for (int64_t z = 0; z < z_num; ++z) {
            const int tn =
#ifdef _OPENMP
                omp_get_thread_num();
#else
                0;
#endif
            RingBuf& b = bufs[tn];
            int64_t zs = std::max<int64_t>(z - hz, 0), ze = std::min<int64_t>(z + hz + 1, z_num);
            for (int64_t iz = zs; iz < ze; ++iz)
                for (int64_t ix = 0; ix <= hx; ++ix) update(b, iz, ix);
            for (int64_t iz = z; iz < hz; ++iz) bc_z(b, iz, true);
            for (int64_t iz = z_num - hz; iz <= z; ++iz) bc_z(b, iz + 2 * hz, false);
            for (int64_t ix = 0; ix < hx; ++ix) bc_x(b, ix, true);
            run(b, z, 0);
            for (int64_t x = 1; x < x_num; ++x) {
                if (x < x_num - hx) {
                    for (int64_t iz = zs; iz < ze; ++iz) update(b, iz, x + hx);
                    for (int64_t iz = z; iz < hz; ++iz) bc_z(b, iz, true);
                    for (int64_t iz = z_num - hz; iz <= z; ++iz) bc_z(b, iz + 2 * hz, false);
                } else {
                    bc_x(b, x + 2 * hx, false);
                }
                run(b, z, x);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Face-neighbour filters: APRNumerics::face_neighbour_filter                 */
/* (APRNumerics.hpp:305-355) with the random iterator's neighbour search      */
/* (APRIterator.hpp:283-394, RandomAccess.hpp:92-170, RandomIterator.hpp:     */
/* 76-91).  For each face (+d, then -d) of a particle at (l, z, x, y):        */
/*   the same-level cell, else (l > l_min) the parent-level cell containing   */
/*   the face neighbour, else (l < l_max) the four children on the face, each */
/*   counted when it is a particle -- provided the first of them is one;      */
/* out = in * f[1], then += f[2] (face +d) and f[0] (face -d) times the mean  */
/* of the neighbours, or times the particle's own value when it has none.    */
/* Coordinates are unsigned 16-bit in the reference: -1 wraps out of range.  */
/* ------------------------------------------------------------------------ */
/* index of the particle at (level, z, x, y), or -1.  The reference's random access resolves
   the four l_max rows under the first parent row, (z, x) in {0, 1}^2, to row (0, 0): their
   particles read the values of row (0, 0) with the same y. */
static inline int64_t find_particle(const GenInfo& g, const Linear& L, int level, int64_t z, int64_t x, int64_t y) {
    if (level < g.l_min || level > g.l_max) return -1;
    if (z < 0 || x < 0 || y < 0 || z >= g.z_num[level] || x >= g.x_num[level] || y >= g.y_num[level]) return -1;
    if (level == g.l_max && z < 2 && x < 2) {
        z = 0;
        x = 0;
    }

    uint64_t b = L.row_begin(level, z, x, g), e = L.row_end(level, z, x, g);
    const uint16_t* yv = L.y_vec.data();
    uint64_t lo = b, hi = e;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if ((int64_t)yv[mid] < y) lo = mid + 1;
        else hi = mid;
    }
    return (lo < e && (int64_t)yv[lo] == y) ? (int64_t)lo : -1;
}

/* the reference's uint16 coordinate arithmetic: a step below 0 wraps to 65535 */
static inline int64_t u16c(int64_t v) { return (int64_t)(uint16_t)v; }

template <typename S>
void face_neighbour_filter(const GenInfo& g, const Linear& L, const S* in, float* out, const float* filter, int dim) {
    static const int DY[6] = {1, -1, 0, 0, 0, 0}, DX[6] = {0, 0, 1, -1, 0, 0}, DZ[6] = {0, 0, 0, 0, 1, -1};
    static const int CO[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
    const float ft[2] = {filter[2], filter[0]};
    for (int level = g.l_min; level <= g.l_max; ++level) {
#pragma omp parallel for schedule(dynamic)
        for (int64_t z = 0; z < g.z_num[level]; ++z)
            for (int64_t x = 0; x < g.x_num[level]; ++x)
                for (uint64_t i = L.row_begin(level, z, x, g); i < L.row_end(level, z, x, g); ++i) {
                    const int64_t y = L.y_vec[i];
                    const float cur = (float)in[i];
                    float o = cur * filter[1];
                    for (int k = 0; k < 2; ++k) {
                        const int f = 2 * dim + k;
                        float sum = 0, cnt = 0;
                        int64_t n = find_particle(g, L, level, u16c(z + DZ[f]), u16c(x + DX[f]), u16c(y + DY[f]));
                        if (n >= 0) {
                            sum += (float)in[n];
                            cnt++;
                        } else {
                            n = (level > g.l_min) ? find_particle(g, L, level - 1, u16c(z + DZ[f]) / 2, u16c(x + DX[f]) / 2,
                                                                  u16c(y + DY[f]) / 2)
                                                  : -1;
                            if (n >= 0) {
                                sum += (float)in[n];
                                cnt++;
                            } else if (level < g.l_max) {
                                /* children on the face: the face axis at 2 (c + d) + (d < 0), the other two
                                   axes 2 c + offsets (CO), the first of them decides */
                                for (int c = 0; c < 4; ++c) {
                                    int64_t cy = u16c(u16c(y + DY[f]) * 2 + (DY[f] < 0)), cx = u16c(u16c(x + DX[f]) * 2 + (DX[f] < 0)),
                                            cz = u16c(u16c(z + DZ[f]) * 2 + (DZ[f] < 0));
                                    if (dim == 0) { cx += CO[c][0]; cz += CO[c][1]; }
                                    else if (dim == 1) { cy += CO[c][0]; cz += CO[c][1]; }
                                    else { cy += CO[c][0]; cx += CO[c][1]; }
                                    const int64_t m = find_particle(g, L, level + 1, u16c(cz), u16c(cx), u16c(cy));
                                    if (m < 0) {
                                        if (c == 0) break;
                                        continue;
                                    }
                                    sum += (float)in[m];
                                    cnt++;
                                }
                            }
                        }
                        if (cnt > 0) o += ft[k] * sum / cnt;
                        else o += ft[k] * cur;
                    }
                    out[i] = o;
                }
    }
}

/* seperable_face_neighbour_filter (APRNumerics.hpp:286-302): y, x, z per repeat */
template <typename S>
void seperable_face_neighbour_filter(const GenInfo& g, const Linear& L, const S* in, float* out, const float* filter, int repeats) {
    const size_t P = L.total;
    std::vector<float> tmp(P), o(P, 0.0f);
    for (size_t i = 0; i < P; ++i) tmp[i] = (float)in[i];
    for (int r = 0; r < repeats; ++r) {
        face_neighbour_filter(g, L, tmp.data(), o.data(), filter, 0);
        face_neighbour_filter(g, L, o.data(), tmp.data(), filter, 1);
        face_neighbour_filter(g, L, tmp.data(), o.data(), filter, 2);
        std::swap(o, tmp);
    }
    std::swap(o, tmp);
    std::copy(o.begin(), o.end(), out);
}

/* APRNumerics::richardson_lucy (APRNumerics.hpp:549-593).  Optional stopping
   rule (not in the reference): stop when ||u_k+1 - u_k||_2 / ||u_k||_2 < tol. */
template <typename InT>
int richardson_lucy(const GenInfo& g, const Linear& L, const Tree& t, const InT* in, float* u, const Mesh<float>& psf, int niter,
                    bool use_ds, bool normalize, bool resume, double tol, double* last_rel) {
    Mesh<float> flipped = psf;
    for (size_t i = 0; i < psf.v.size(); ++i) flipped.v[i] = psf.v[psf.v.size() - 1 - i];
    const int nst = use_ds ? g.l_max - g.l_min : 1;
    std::vector<Mesh<float>> sv, fv;
    downsampled_stencils(psf, sv, nst, normalize);
    downsampled_stencils(flipped, fv, nst, normalize);
    const size_t P = L.total;
    if (!resume) std::fill(u, u + P, 1.0f);
    std::vector<float> blur(P), est(P), tr;
    int it = 0;
    double rel = 0;
    for (it = 0; it < niter; ++it) {
        fill_tree_mean(g, L, t, u, tr);
        convolve_pencil(g, L, t, fv, u, tr.data(), blur.data(), true);
        for (size_t i = 0; i < P; ++i) blur[i] = (float)in[i] / blur[i];
        fill_tree_mean(g, L, t, blur.data(), tr);
        convolve_pencil(g, L, t, sv, blur.data(), tr.data(), est.data(), true);
        double num = 0, den = 0;
        for (size_t i = 0; i < P; ++i) {
            const float nu = u[i] * est[i];
            const double d = (double)nu - (double)u[i];
            num += d * d;
            den += (double)u[i] * (double)u[i];
            u[i] = nu;
        }
        rel = den > 0 ? std::sqrt(num / den) : 0.0;
        if (tol > 0 && rel < tol) {
            ++it;
            break;
        }
    }
    if (last_rel) *last_rel = rel;
    return it;
}

/* APRNumerics::gradient_cfd (APRNumerics.hpp:163-184): the central difference
   [-1, 0, 1] / (2 delta) along dim (0 y, 1 x, 2 z), divided by 2^d at level delta d
   (APRStencil::get_rescaled_stencils, APRStencil.hpp:135-152), convolve_pencil with reflected
   ends and the input's tree means */
template <typename PartT>
void gradient_cfd(const GenInfo& g, const Linear& L, const Tree& t, const PartT* in, float* out, int dim, float delta) {
    Mesh<float> s(dim == 0 ? 3 : 1, dim == 1 ? 3 : 1, dim == 2 ? 3 : 1);
    s.v[0] = -1.0f / (2 * delta);
    s.v[1] = 0.0f;
    s.v[2] = 1.0f / (2 * delta);
    const int nl = std::max(g.l_max - g.l_min, 1);
    std::vector<Mesh<float>> st((size_t)nl, s);
    for (int d = 0; d < nl; ++d) {
        const float step = (float)std::pow(2.0, (double)d);
        for (auto& w : st[(size_t)d].v) w = w / step;
    }
    std::vector<float> tr;
    fill_tree_mean(g, L, t, in, tr);
    convolve_pencil(g, L, t, st, in, tr.data(), out, true);
}

/* APRNumerics::richardson_lucy_tv (APRNumerics.hpp:489-525, :549-572): Richardson-Lucy whose
   update divides by 1 - reg * div(grad u / |grad u|); the divergence of the normalised
   gradient as div_norm_grad (:431-486) with unit pixel sizes: central differences of u along
   each axis present, each normalised by sqrtf(g_z^2 + g_x^2 + g_y^2) where that exceeds 1e-6,
   then the differences of the normalised components summed y, + x, + z */
template <typename InT>
void richardson_lucy_tv(const GenInfo& g, const Linear& L, const Tree& t, const InT* in, float* u, const Mesh<float>& psf, int niter,
                        float reg, bool use_ds, bool normalize, bool resume) {
    Mesh<float> flipped = psf;
    for (size_t i = 0; i < psf.v.size(); ++i) flipped.v[i] = psf.v[psf.v.size() - 1 - i];
    const int nst = use_ds ? g.l_max - g.l_min : 1;
    std::vector<Mesh<float>> sv, fv;
    downsampled_stencils(psf, sv, nst, normalize);
    downsampled_stencils(flipped, fv, nst, normalize);
    const size_t P = L.total;
    if (!resume) std::fill(u, u + P, 1.0f);
    std::vector<float> blur(P), est(P), gx(P), gy(P), gz(P), div(P), tmp(P), tr;
    for (int it = 0; it < niter; ++it) {
        fill_tree_mean(g, L, t, u, tr);
        convolve_pencil(g, L, t, fv, u, tr.data(), blur.data(), true);
        for (size_t i = 0; i < P; ++i) blur[i] = (float)in[i] / blur[i];
        fill_tree_mean(g, L, t, blur.data(), tr);
        convolve_pencil(g, L, t, sv, blur.data(), tr.data(), est.data(), true);
        gradient_cfd(g, L, t, u, gy.data(), 0, 1.0f);
        if (g.org[1] > 1) gradient_cfd(g, L, t, u, gx.data(), 1, 1.0f);
        else std::fill(gx.begin(), gx.end(), 0.0f);
        if (g.org[2] > 1) gradient_cfd(g, L, t, u, gz.data(), 2, 1.0f);
        else std::fill(gz.begin(), gz.end(), 0.0f);
        for (size_t i = 0; i < P; ++i) {
            const float m = std::sqrt(gz[i] * gz[i] + gx[i] * gx[i] + gy[i] * gy[i]);
            if (m > 1e-6) {
                gz[i] /= m;
                gx[i] /= m;
                gy[i] /= m;
            }
        }
        gradient_cfd(g, L, t, gy.data(), div.data(), 0, 1.0f);
        if (g.x_num[g.l_max] > 1) {
            gradient_cfd(g, L, t, gx.data(), tmp.data(), 1, 1.0f);
            for (size_t i = 0; i < P; ++i) div[i] = tmp[i] + div[i];
        }
        if (g.z_num[g.l_max] > 1) {
            gradient_cfd(g, L, t, gz.data(), tmp.data(), 2, 1.0f);
            for (size_t i = 0; i < P; ++i) div[i] = tmp[i] + div[i];
        }
        for (size_t i = 0; i < P; ++i) u[i] = u[i] * est[i] / (1.0f - reg * div[i]);
    }
}

/* APRReconstruction::reconstruct_constant (APRReconstruction.hpp:80-120) and the
   full-volume patch form with level_delta <= 0 (:297-394): the image at level
   L = l_max + level_delta; particles of level l <= L fill blocks of 2^(L-l) (clipped
   at the level's edge), and for L < l_max the tree means of level L overwrite the
   cells where finer particles live.  Values convert to the image type like a C++
   assignment on x86-64 (float -> u16 through int32).                          */
template <typename ImgT> struct ImgCvt { static ImgT of(float v) { return (ImgT)v; } };
template <> struct ImgCvt<uint16_t> { static uint16_t of(float v) { return f2u16(v); } };

template <typename PartT, typename ImgT>
void reconstruct_constant(const GenInfo& g, const Linear& L, const Tree* t, const PartT* parts, const float* tree, int level_delta,
                          Mesh<ImgT>& img) {
    const int lv = g.l_max + level_delta;
    const int64_t yn = g.y_num[lv], xn = g.x_num[lv], zn = g.z_num[lv];
    img.resize(yn, xn, zn, ImgT(0));
    for (int64_t z = 0; z < zn; ++z)
        for (int64_t x = 0; x < xn; ++x)
            for (int level = lv; level >= g.l_min; --level) {
                const int64_t step = (int64_t)1 << (lv - level);
                const int64_t xl = x / step, zl = z / step;
                for (uint64_t i = L.row_begin(level, zl, xl, g); i < L.row_end(level, zl, xl, g); ++i) {
                    const int64_t y0 = (int64_t)L.y_vec[i] * step;
                    const int64_t y1 = std::min(y0 + step, yn);
                    const ImgT v = ImgCvt<ImgT>::of((float)parts[i]);
                    for (int64_t y = y0; y < y1; ++y) img.at(y, x, z) = v;
                }
            }
    if (lv < g.l_max) {
        const GenInfo& ti = t->info;
        for (int64_t z = 0; z < zn; ++z)
            for (int64_t x = 0; x < xn; ++x)
                for (uint64_t i = t->lin.row_begin(lv, z, x, ti); i < t->lin.row_end(lv, z, x, ti); ++i)
                    img.at(t->lin.y_vec[i], x, z) = ImgCvt<ImgT>::of(tree[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* APRReconstruction::reconstruct_smooth (APRReconstruction.hpp:270-295): the  */
/* constant reconstruction, then per axis a running-sum mean whose half-width */
/* at a pixel is min(floor(2^(l_max - level) / scale), offset_max) with       */
/* offset_max = min(20, n/2 - 1) (calc_sat_adaptive_y/x/z, :528-876).  The    */
/* arithmetic is the reference's: cumulative sums in float, some scale        */
/* factors in double, every store converted to the image type T as a C++     */
/* assignment on x86-64 (float / double -> u16 through int32), the half-width */
/* as T then as float (y) or int64 (x, z).  As in the reference, the z pass's */
/* low-end boundary test compares the x index (not z) with the half-width.    */
/* ------------------------------------------------------------------------ */
template <typename T> struct SmCvt;
template <> struct SmCvt<float> {
    static float f(float v) { return v; }
    static float d(double v) { return (float)v; }
    static float off(double v) { return (float)v; }
};
template <> struct SmCvt<uint16_t> {
    static uint16_t f(float v) { return f2u16(v); }
    static uint16_t d(double v) { return d2u16(v); }
    static uint16_t off(double v) { return d2u16(v); }
};

template <typename T>
static inline T sm_offset(uint16_t level, float scale, int om, int d_max) {
    const T a = SmCvt<T>::off(std::floor(std::pow(2.0, (double)(d_max - (int)level)) / (double)scale));
    const T b = (T)om;
    return std::min(a, b);
}

template <typename T>
void sat_adaptive_y(Mesh<T>& in, const Mesh<uint16_t>& kimg, float scale, int om_in, int d_max) {
    const int64_t yn = in.y_num, xn = in.x_num, zn = in.z_num;
    const int om = std::min(om_in, (int)(yn / 2 - 1));
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < zn; ++j) {
        std::vector<float> tv((size_t)yn), ov((size_t)yn);
        for (int64_t i = 0; i < xn; ++i) {
            T* m = in.v.data() + (size_t)((j * xn + i) * yn);
            const uint16_t* kl = kimg.v.data() + (size_t)((j * xn + i) * yn);
            for (int64_t k = 0; k < yn; ++k) ov[k] = (float)sm_offset<T>(kl[k], scale, om, d_max);
            float temp = 0;
            for (int64_t k = 0; k < yn; ++k) {
                temp += (float)m[k];
                tv[k] = temp;
            }
            m[0] = 0;
            for (int64_t k = 1; k <= om + 1; ++k) {
                const float div = 2 * ov[k] + 1, off = ov[k];
                if ((float)k <= off + 1) m[k] = SmCvt<T>::f(-tv[0] / div);
            }
            for (int64_t k = 0; k < yn; ++k) {
                const float div = 2 * ov[k] + 1, off = ov[k];
                if ((float)k >= off + 1) m[k] = SmCvt<T>::f(-tv[(size_t)(k - (int64_t)off - 1)] / div);
            }
            for (int64_t k = 0; k < yn; ++k) {
                const float div = 2 * ov[k] + 1, off = ov[k];
                if ((float)k < (float)yn - off) m[k] = SmCvt<T>::f((float)m[k] + tv[(size_t)(k + (int64_t)off)] / div);
            }
            for (int64_t k = yn - om; k < yn; ++k) {
                const float div = 2 * ov[k] + 1, off = ov[k];
                if ((float)k >= (float)yn - off) {
                    const float counter = (float)k - ((float)yn - off) + 1;
                    m[k] = SmCvt<T>::f((float)m[k] * div);
                    m[k] = SmCvt<T>::f((float)m[k] + tv[(size_t)(yn - 1)]);
                    m[k] = SmCvt<T>::d((double)m[k] * (1.0 / (double)(div - counter)));
                }
            }
            for (int64_t k = 1; k < om + 1; ++k) {
                const float div = 2 * ov[k] + 1, off = ov[k];
                if ((float)k < off + 1) m[k] = SmCvt<T>::d((double)m[k] * ((double)div / (1.0 * k + (double)off)));
            }
            {
                const float div = 2 * ov[0] + 1, off = ov[0];
                m[0] = SmCvt<T>::f((float)m[0] * (div / (off + 1)));
            }
        }
    }
}

/* x (axis 1) or z (axis 2) pass: a ring of 2 offset_max + 2 cumulative-sum rows */
template <typename T>
void sat_adaptive_outer(Mesh<T>& in, const Mesh<uint16_t>& kimg, float scale, int om_in, int d_max, int axis) {
    const int64_t yn = in.y_num, xn = in.x_num, zn = in.z_num;
    const int64_t n = axis == 1 ? xn : zn;
    const int64_t om = std::min(om_in, (int)(n / 2 - 1));
    const int64_t R = 2 * om + 2;
    const int64_t outer = axis == 1 ? zn : xn;
    const int64_t stride = axis == 1 ? yn : xn * yn;
#pragma omp parallel for schedule(static)
    for (int64_t o = 0; o < outer; ++o) {
        std::vector<float> tv((size_t)(yn * R), 0.0f);
        const int64_t base = axis == 1 ? o * xn * yn : o * yn;
        T* m = in.v.data() + base;
        const uint16_t* kl = kimg.v.data() + base;
        auto M = [&](int64_t i, int64_t k) -> T& { return m[i * stride + k]; };
        auto K = [&](int64_t i, int64_t k) { return kl[i * stride + k]; };
        for (int64_t k = 0; k < yn; ++k) tv[(size_t)k] = (float)M(0, k);
        for (int64_t i = 1; i < 2 * om + 1; ++i)
            for (int64_t k = 0; k < yn; ++k) tv[(size_t)(i * yn + k)] = (float)M(i, k) + tv[(size_t)((i - 1) * yn + k)];
        for (int64_t i = 0; i < om + 1; ++i)
            for (int64_t k = 0; k < yn; ++k) {
                const int64_t off = (int64_t)sm_offset<T>(K(i, k), scale, (int)om, d_max);
                /* the z pass tests the x index here (APRReconstruction.hpp:821) */
                const int64_t t = axis == 1 ? i : o;
                if (t < off + 1) M(i, k) = SmCvt<T>::f(tv[(size_t)((i + off) * yn + k)] / (float)(i + off + 1));
            }
        for (int64_t i = 1; i < n; ++i)
            for (int64_t k = 0; k < yn; ++k) {
                const int64_t off = (int64_t)sm_offset<T>(K(i, k), scale, (int)om, d_max);
                if ((i >= om + 1) & (i < n - om)) {
                    const int64_t im = (i + om) % R, pm = (i + om - 1) % R;
                    tv[(size_t)(im * yn + k)] = (float)M(i + om, k) + tv[(size_t)(pm * yn + k)];
                }
                if ((i >= off + 1) & (i < n - off)) {
                    const int64_t fm = (i + off) % R, bm = (i - off - 1) % R;
                    M(i, k) = SmCvt<T>::f((tv[(size_t)(fm * yn + k)] - tv[(size_t)(bm * yn + k)]) / (float)(2 * off + 1));
                }
            }
        for (int64_t i = n - om; i < n; ++i)
            for (int64_t k = 0; k < yn; ++k) {
                const int64_t off = (int64_t)sm_offset<T>(K(i, k), scale, (int)om, d_max);
                if (i >= n - off) {
                    const int64_t bm = (i - off - 1) % R, fm = (n - 1) % R;
                    M(i, k) = SmCvt<T>::f((tv[(size_t)(fm * yn + k)] - tv[(size_t)(bm * yn + k)]) / (float)(n - i + off));
                }
            }
    }
}

/* the level of the particle covering each pixel (reconstruct_level, APRReconstruction.hpp:131-171) */
static inline void level_image(const GenInfo& g, const Linear& L, Mesh<uint16_t>& img) {
    img.resize(g.y_num[g.l_max], g.x_num[g.l_max], g.z_num[g.l_max], 0);
    for (int level = g.l_min; level <= g.l_max; ++level) {
        const int64_t s = (int64_t)1 << (g.l_max - level);
        for (int64_t z = 0; z < g.z_num[level]; ++z)
            for (int64_t x = 0; x < g.x_num[level]; ++x)
                for (uint64_t i = L.row_begin(level, z, x, g); i < L.row_end(level, z, x, g); ++i)
                    for (int64_t zz = z * s; zz < std::min((z + 1) * s, img.z_num); ++zz)
                        for (int64_t xx = x * s; xx < std::min((x + 1) * s, img.x_num); ++xx)
                            for (int64_t yy = L.y_vec[i] * s; yy < std::min((L.y_vec[i] + 1) * s, img.y_num); ++yy)
                                img.at(yy, xx, zz) = (uint16_t)level;
    }
}

template <typename PartT, typename ImgT>
void reconstruct_smooth(const GenInfo& g, const Linear& L, const PartT* parts, const float* scale_d, Mesh<ImgT>& img) {
    reconstruct_constant<PartT, ImgT>(g, L, nullptr, parts, nullptr, 0, img);
    Mesh<uint16_t> k;
    level_image(g, L, k);
    const int om = 20;
    if (img.y_num > 1) sat_adaptive_y(img, k, scale_d[0], om, g.l_max);
    if (img.x_num > 1) sat_adaptive_outer(img, k, scale_d[1], om, g.l_max, 1);
    if (img.z_num > 1) sat_adaptive_outer(img, k, scale_d[2], om, g.l_max, 2);
}

/* APRStencil::create_gaussian_filter (APRStencil.hpp:173-228) */
static inline void gaussian_stencil(const float* sigma, const int* size, bool normalize, Mesh<float>& st) {
    st.resize(size[0], size[1], size[2], 0.0f);
    std::vector<float> g[3];
    for (int d = 0; d < 3; ++d) {
        g[d].assign((size_t)size[d], 1.0f);
        if (size[d] > 1)
            for (int i = -size[d] / 2, c = 0; i <= size[d] / 2; ++i, ++c)
                g[d][(size_t)c] = expf(-i * i / (2 * sigma[d] * sigma[d])) / (sigma[d] * sqrtf(2 * M_PI));
    }
    float sum = 0;
    for (int i = 0; i < size[0]; ++i)
        for (int j = 0; j < size[1]; ++j)
            for (int k = 0; k < size[2]; ++k) {
                st.at(i, j, k) = g[0][(size_t)i] * g[1][(size_t)j] * g[2][(size_t)k];
                sum += st.at(i, j, k);
            }
    if (normalize)
        for (auto& v : st.v) v /= sum;
}

}  // namespace oracle

#endif
