/*
 * k_grad_lis.hip -- downsampled gradient, inverse B-spline, local intensity
 * scale and level computation of the conversion pipeline (gfx950).
 *
 *  - k_grad_ds: one lane per downsampled cell; gradient magnitude of its
 *    2x2x2 voxels (ComputeGradient.hpp:975-1051, double squares/sqrt like the
 *    reference) fused with the 2x2x2 mean downsample (ComputeGradient.hpp:95-98).
 *  - inverse B-spline: all active axes in one z-marching sweep, out of place
 *    (ComputeGradient.hpp:784-964).
 *  - local intensity scale (LocalIntensityScale.hpp:26-494): the six
 *    separable running-sum means over the reflect-padded array.  The y mean
 *    is a per-line sequential prefix sum, so lines are staged through an LDS
 *    tile (64 lines per wave) and each lane walks its line with a register
 *    delay line of 2*offset+2 prefix sums; the x / z means run one lane per
 *    y with the delay line along the outer axis.  Every float operation keeps
 *    the reference's order and precision (float vs double sub-expressions).
 *  - k_levels: applyParameters + computeLevels fused (APRConverter.hpp:249-297,
 *    LocalParticleCellSet.hpp:128-171).
 */
#include <algorithm>
#include <cmath>

#include "aprh_internal.hpp"
#include "tiles.hpp"

namespace aprh {

namespace {

__device__ __forceinline__ uint16_t d2u16(double v) {
    if (!(v > -2147483648.0 && v < 2147483648.0)) return 0;
    return (uint16_t)(int32_t)v;
}

__device__ __forceinline__ uint32_t f2u32(float v) {
    if (!(v > -9.2233720368547758e18f && v < 9.2233720368547758e18f)) return 0u;
    return (uint32_t)(int64_t)v;
}

__device__ __forceinline__ double sq_term(int d, float h2) {
    const double q = (double)((float)d / h2);
    return q * q;
}

/* difference of two samples as the reference computes it: int for uint16 (promoted, then
   converted to float for the division), float for float images */
__device__ __forceinline__ float gdiff(uint16_t a, uint16_t b) { return (float)((int)a - (int)b); }
__device__ __forceinline__ float gdiff(float a, float b) { return a - b; }
__device__ __forceinline__ double gterm(float d, float h2) {
    const double q = (double)(d / h2);
    return q * q;
}
__device__ __forceinline__ void gstore(double v, uint16_t& out) { out = d2u16(v); }
__device__ __forceinline__ void gstore(double v, float& out) { out = (float)v; }

template <typename T>
__device__ __forceinline__ void grad_ds_body(const T* __restrict__ img, T* grad, float* mean, int64_t y_num, int64_t x_num,
                                             int64_t z_num, int64_t yd_num, float h2x, float h2y, float h2z, int accum, int64_t bx,
                                             int64_t xd, int64_t zd, int64_t xd_num) {
    const int64_t yd = bx * blockDim.x + threadIdx.x;
    if (yd >= yd_num) return;
    const int64_t xy = x_num * y_num;
    T g = accum ? grad[(zd * xd_num + xd) * yd_num + yd] : (T)0;
    const int64_t y0 = 2 * yd, y1 = min(2 * yd + 1, y_num - 1);
    for (int64_t z = 2 * zd; z <= min(2 * zd + 1, z_num - 1); ++z) {
        const int64_t zm = z > 0 ? z - 1 : 0, zp = min(z + 1, z_num - 1);
        for (int64_t x = 2 * xd; x <= min(2 * xd + 1, x_num - 1); ++x) {
            const int64_t xm = x > 0 ? x - 1 : 0, xp = min(x + 1, x_num - 1);
            const T* cen = img + z * xy + x * y_num;
            const T* left = img + z * xy + xm * y_num;
            const T* right = img + z * xy + xp * y_num;
            const T* up = img + zm * xy + x * y_num;
            const T* down = img + zp * xy + x * y_num;
            for (int w = 0; w < 2; ++w) {
                const int64_t y = w ? y1 : y0;
                T t;
                if (y_num >= 2) {
                    float dyv;
                    if (y == 0) dyv = gdiff(cen[1], cen[0]);
                    else if (y == y_num - 1) dyv = gdiff(cen[y_num - 1], cen[y_num - 2]);
                    else dyv = gdiff(cen[y + 1], cen[y - 1]);
                    const double s = gterm(gdiff(right[y], left[y]), h2x) + gterm(gdiff(down[y], up[y]), h2z) + gterm(dyv, h2y);
                    gstore(sqrt(s), t);
                } else {
                    t = (T)0;
                }
                g = t > g ? t : g;
            }
        }
    }
    grad[(zd * xd_num + xd) * yd_num + yd] = g;
    if (mean) {
        /* downsample(sum, /8.0): clamped indices, fixed reduce order (PixelData.hpp:944-962) */
        const int64_t shx = min(2 * xd + 1, x_num - 1), shz = min(2 * zd + 1, z_num - 1), shy = y1;
        const int64_t z0 = 2 * zd, x0 = 2 * xd;
        float s = (float)img[z0 * xy + x0 * y_num + y0];
        s = s + (float)img[z0 * xy + x0 * y_num + shy];
        s = s + (float)img[z0 * xy + shx * y_num + y0];
        s = s + (float)img[z0 * xy + shx * y_num + shy];
        s = s + (float)img[shz * xy + x0 * y_num + y0];
        s = s + (float)img[shz * xy + x0 * y_num + shy];
        s = s + (float)img[shz * xy + shx * y_num + y0];
        s = s + (float)img[shz * xy + shx * y_num + shy];
        mean[(zd * xd_num + xd) * yd_num + yd] = (float)((double)s / 8.0);
    }
}


/* Vector form of k_grad_ds for y_num % 8 == 0: a lane owns 8 consecutive y of
 * one downsampled (x, z) column pair, i.e. 4 downsampled cells.  It loads the
 * 12 full-resolution rows its 2x2 block and their x/z neighbours touch as
 * 16-byte vectors (plus the y-1 / y+8 samples of the centre rows), computes
 * the 32 gradient magnitudes, and writes 4 maxima and 4 means.
 * Unit spacing (hx = hy = hz = 1) uses the exact integer form of the
 * reference's double arithmetic: each squared term is d^2/4, the sum k/4 is
 * exact, and trunc(sqrt(k/4)) == floor(sqrt(k)) >> 1.  Both the square root and the
 * truncation are monotone, so a cell's maximum magnitude is that of its largest k: the
 * lanes take the integer maximum of k over the cell's voxels and one square root per cell
 * (the same holds for the double form of other spacings). */
/* floor(sqrt(k)) for k < 2^35: the hardware square root of the float-converted k carries a
 * relative error below 2^-22 (one rounding of k, one ulp of v_sqrt_f32), i.e. an absolute
 * error below 0.05 for sqrt(k) < 2^17.5, so its truncation is within one of the answer and
 * one exact 64-bit correction step finishes it. */
__device__ __forceinline__ uint32_t isqrt_k(uint64_t k) {
    uint32_t s = (uint32_t)__builtin_amdgcn_sqrtf((float)k);
    if ((uint64_t)s * s > k) --s;
    else if ((uint64_t)(s + 1) * (s + 1) <= k) ++s;
    return s;
}

__device__ __forceinline__ int u16at(const uint4& v, int i) {
    const uint32_t w = (i < 2) ? v.x : (i < 4) ? v.y : (i < 6) ? v.z : v.w;
    return (int)((i & 1) ? (w >> 16) : (w & 0xFFFFu));
}

/* One lane owns 8 consecutive y of one downsampled x column pair and marches over a run of
 * GZC downsampled z planes.  Step zd needs the full-resolution planes 2zd-1 .. 2zd+2; the
 * rows (x0, x1) of plane 2zd-1 and of plane 2zd are the previous step's plane-(2zd-1) centre
 * rows and its down rows, so a step loads 8 rows instead of 12. */
constexpr int GZC = 8;
template <bool UNIT>
__device__ __forceinline__ void grad_ds_vec_body(const uint16_t* __restrict__ img, uint16_t* grad, float* mean, int64_t y_num,
                                                 int64_t x_num, int64_t z_num, int64_t yd_num, float h2x, float h2y, float h2z,
                                                 int accum, int64_t yg, int64_t xd, int64_t zc, int64_t xd_num, int64_t zd_num) {
    /* yg: group of 8 y */
    if (yg * 8 >= y_num) return;
    const int64_t xy = x_num * y_num;
    const int64_t y0 = yg * 8;
    const int64_t x0 = 2 * xd;
    const bool hx1 = x0 + 1 < x_num;
    const int64_t x1 = hx1 ? x0 + 1 : x0;
    auto row = [&](int64_t x, int64_t z) -> const uint16_t* { return img + z * xy + x * y_num + y0; };
    auto ld = [&](int64_t x, int64_t z) -> uint4 { return *(const uint4*)row(x, z); };
    const int64_t xl0 = x0 > 0 ? x0 - 1 : 0, xr1 = min(x1 + 1, x_num - 1);
    /* y neighbours beyond the vector (replicate at the line ends) */
    auto edge = [&](int64_t x, int64_t z, int which) -> int {
        const uint16_t* p = img + z * xy + x * y_num;
        if (which == 0) return y0 > 0 ? (int)p[y0 - 1] : -1;
        return y0 + 8 < y_num ? (int)p[y0 + 8] : -1;
    };
    const int64_t zd_begin = zc * GZC, zd_end = min(zd_num, zd_begin + GZC);
    uint4 keep_u0 = make_uint4(0, 0, 0, 0), keep_u1 = keep_u0, keep_c0 = keep_u0, keep_c1 = keep_u0;
    int64_t keep_uz = -1, keep_cz = -1;   /* planes the kept rows belong to */
    for (int64_t zd = zd_begin; zd < zd_end; ++zd) {
        const int64_t z0 = 2 * zd;
        const bool hz1 = z0 + 1 < z_num;
        const int64_t z1 = hz1 ? z0 + 1 : z0;
        const int64_t zu0 = z0 > 0 ? z0 - 1 : 0, zd1 = min(z1 + 1, z_num - 1);
        /* centre rows c[a][b]: x = a ? x1 : x0, z = b ? z1 : z0 */
        const bool kc = keep_cz == z0, ku = keep_uz == zu0;
        const uint4 c00 = kc ? keep_c0 : ld(x0, z0), c10 = kc ? keep_c1 : ld(x1, z0);
        const uint4 c01 = ld(x0, z1), c11 = ld(x1, z1);
        const uint4 l0 = ld(xl0, z0), l1 = ld(xl0, z1), r0 = ld(xr1, z0), r1 = ld(xr1, z1);
        const uint4 u0 = ku ? keep_u0 : ld(x0, zu0), u1 = ku ? keep_u1 : ld(x1, zu0);
        const uint4 d0 = ld(x0, zd1), d1 = ld(x1, zd1);
        const int e00m = edge(x0, z0, 0), e00p = edge(x0, z0, 1), e10m = edge(x1, z0, 0), e10p = edge(x1, z0, 1);
        const int e01m = edge(x0, z1, 0), e01p = edge(x0, z1, 1), e11m = edge(x1, z1, 0), e11p = edge(x1, z1, 1);
        /* per cell (y pair): the largest squared-difference sum (unit spacing) or the largest
           double magnitude before the square root */
        uint64_t kmax[4] = {0, 0, 0, 0};
        double qmax[4] = {0.0, 0.0, 0.0, 0.0};
        auto one_row = [&](const uint4& c, const uint4& lft, const uint4& rgt, const uint4& up, const uint4& dn, int em, int ep) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int64_t y = y0 + i;
                int dyv;
                /* y == 0: c[1]-c[0]; y == Y-1: c[Y-1]-c[Y-2]; else c[y+1]-c[y-1] */
                const int cm = (i > 0) ? u16at(c, i - 1) : em;
                const int cp = (i < 7) ? u16at(c, i + 1) : ep;
                if (y == 0) dyv = u16at(c, 1) - u16at(c, 0);
                else if (y == y_num - 1) dyv = u16at(c, i) - cm;
                else dyv = cp - cm;
                const int dxv = u16at(rgt, i) - u16at(lft, i);
                const int dzv = u16at(dn, i) - u16at(up, i);
                if (UNIT) {
                    /* |d| <= 65535: the square of (uint32)d modulo 2^32 is d^2 exactly */
                    const uint32_t ux = (uint32_t)dxv, uy = (uint32_t)dyv, uz = (uint32_t)dzv;
                    const uint64_t k = (uint64_t)(ux * ux) + (uint64_t)(uy * uy) + (uint64_t)(uz * uz);
                    kmax[i >> 1] = max(kmax[i >> 1], k);
                } else {
                    const double q = sq_term(dxv, h2x) + sq_term(dzv, h2z) + sq_term(dyv, h2y);
                    qmax[i >> 1] = fmax(qmax[i >> 1], q);
                }
            }
        };
        /* rows (x0,z0), (x1,z0), (x0,z1), (x1,z1); left/right neighbours of x1 are x0 / xr1 etc. */
        one_row(c00, l0, hx1 ? c10 : r0, u0, hz1 ? c01 : d0, e00m, e00p);
        if (hx1) one_row(c10, c00, r0, u1, hz1 ? c11 : d1, e10m, e10p);
        if (hz1) {
            one_row(c01, l1, hx1 ? c11 : r1, c00, d0, e01m, e01p);
            if (hx1) one_row(c11, c01, r1, c10, d1, e11m, e11p);
        }
        uint32_t gmax[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) gmax[j] = UNIT ? isqrt_k(kmax[j]) >> 1 : (uint32_t)d2u16(sqrt(qmax[j]));
        const int64_t o = (zd * xd_num + xd) * yd_num + yg * 4;
        if (accum) {
#pragma unroll
            for (int j = 0; j < 4; ++j) gmax[j] = max(gmax[j], (uint32_t)grad[o + j]);
        }
        uint2 gw;
        gw.x = gmax[0] | (gmax[1] << 16);
        gw.y = gmax[2] | (gmax[3] << 16);
        *(uint2*)(grad + o) = gw;
        if (mean) {
            /* downsample(sum, /8.0) with clamped x/z (PixelData.hpp:944-962); exact integer sums */
            float4 m;
            float* mp = &m.x;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int a = 2 * j, b = 2 * j + 1;
                const int sum = u16at(c00, a) + u16at(c00, b) + u16at(c10, a) + u16at(c10, b) + u16at(c01, a) + u16at(c01, b) +
                                u16at(c11, a) + u16at(c11, b);
                mp[j] = (float)sum * 0.125f;   /* sum < 2^24: exact, equal to the double division */
            }
            *(float4*)(mean + o) = m;
        }
        /* the next step's up rows are this step's (x0, x1) rows of plane z1, its first centre
           rows this step's down rows */
        keep_u0 = c01;
        keep_u1 = c11;
        keep_uz = z1;
        keep_c0 = d0;
        keep_c1 = d1;
        keep_cz = zd1;
    }
}

/* ---------------- inverse cubic B-spline, [1 4 1]/6 (ComputeGradient.hpp:784-964) ---------------- */
/* All active passes of the inverse B-spline in one sweep.  A block owns a 64 (y) x 16 (x) tile
 * and marches along a run of INV_ZC z planes: each plane is loaded once with a one-sample halo
 * into LDS, the y pass is applied to the 18 rows of the tile (the halo rows' y outputs are the
 * x neighbours), then the x pass, and the x-y result of three consecutive planes gives the z
 * pass of the middle one.  Each value is formed by the same operations in the same order as the
 * reference's separate passes (y: ComputeGradient.hpp:785-830; z: 832-899; x: 901-964; boundary
 * samples with their own two-term forms), so the result is identical to running them one after
 * another. */
constexpr int INV_TY = 64, INV_TX = 16, INV_ZC = 16;

__device__ __forceinline__ float inv_y_at(float l, float c, float r, int64_t k, int64_t n) {
    const float a1 = (float)(1.0 / 6.0), a2 = (float)(4.0 / 6.0), a3 = (float)(1.0 / 6.0);
    float v;
    if (k == 0) {
        v = a2 * c;
        v += (a1 + a3) * r;
    } else if (k == n - 1) {
        v = (a1 + a3) * l;
        v += a2 * c;
    } else {
        v = a1 * l + a2 * c + a3 * r;
    }
    return v;
}
__device__ __forceinline__ float inv_outer_at(float l, float c, float r, int64_t j, int64_t n) {
    const float a1 = (float)(1.0 / 6.0), a2 = (float)(4.0 / 6.0), a3 = (float)(1.0 / 6.0);
    float v;
    if (j == n - 1) {
        v = (a1 + a3) * l;
        v += a2 * c;
    } else if (j == 0) {
        v = a1 * r + a2 * c + a3 * r;
    } else {
        v = a1 * l + a2 * c + a3 * r;
    }
    return v;
}

struct InvFusedArgs {
    const float* in;
    float* out;
    int64_t y, x, z;
    int ay, ax, az;
    int64_t nty, ntx;   /* tiles along y and x */
};

__device__ __forceinline__ void inv_fused_body(const InvFusedArgs& a, int64_t t) {
    constexpr int RW = INV_TY + 2, NR = INV_TX + 2;   /* halo'd row width (y) and row count (x) */
    constexpr int NLD = (RW * NR + 255) / 256;
    __shared__ float s_in[NR][RW + 2];
    __shared__ float s_y[NR][INV_TY + 1];
    const int64_t Y = a.y, X = a.x, Z = a.z;
    const int64_t ty = t % a.nty, r0 = t / a.nty, tx = r0 % a.ntx, tz = r0 / a.ntx;
    const int64_t y0 = ty * INV_TY, x0 = tx * INV_TX, zs = tz * INV_ZC, ze = min(zs + INV_ZC, Z);
    const int ly = threadIdx.x & 63, lx = threadIdx.x >> 6;
    const int64_t yy = y0 + ly;
    /* the plane's halo'd tile, loaded cooperatively into registers */
    auto gload = [&](int64_t zz, float* v) {
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int idx = threadIdx.x + 256 * i;
            const int r = idx / RW, c = idx - r * RW;
            const int64_t gx = x0 - 1 + r, gy = y0 - 1 + c;
            v[i] = (r < NR && gx >= 0 && gx < X && gy >= 0 && gy < Y) ? a.in[(zz * X + gx) * Y + gy] : 0.0f;
        }
    };
    auto lput = [&](const float* v) {
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int idx = threadIdx.x + 256 * i;
            const int r = idx / RW, c = idx - r * RW;
            if (r < NR) s_in[r][c] = v[i];
        }
    };
    /* x-y result of one plane for this thread's four x columns */
    auto plane_xy = [&](float* o) {
        __syncthreads();   /* s_in written */
        if (yy < Y) {
            for (int r = lx; r < NR; r += 4) {
                const int64_t gx = x0 - 1 + r;
                if (gx < 0 || gx >= X) continue;
                const float c = s_in[r][ly + 1];
                s_y[r][ly] = a.ay ? inv_y_at(s_in[r][ly], c, s_in[r][ly + 2], yy, Y) : c;
            }
        }
        __syncthreads();   /* s_y written; s_in free */
#pragma unroll
        for (int cI = 0; cI < 4; ++cI) {
            const int r = 1 + lx * 4 + cI;
            const int64_t gx = x0 - 1 + r;
            float v = 0.0f;
            if (yy < Y && gx < X) v = a.ax ? inv_outer_at(s_y[r - 1][ly], s_y[r][ly], s_y[r + 1][ly], gx, X) : s_y[r][ly];
            o[cI] = v;
        }
    };
    auto store = [&](int64_t zz, const float* v) {
        if (yy >= Y) return;
#pragma unroll
        for (int cI = 0; cI < 4; ++cI) {
            const int64_t gx = x0 + lx * 4 + cI;
            if (gx < X) a.out[(zz * X + gx) * Y + yy] = v[cI];
        }
    };
    const int64_t zf = a.az ? max(zs - 1, (int64_t)0) : zs, zl = a.az ? min(ze, Z - 1) : ze - 1;
    float pm[4] = {0, 0, 0, 0}, pc[4] = {0, 0, 0, 0}, pn[4];
    float cur[NLD], nxt[NLD];
    gload(zf, cur);
    for (int64_t zz = zf; zz <= zl; ++zz) {
        if (zz < zl) gload(zz + 1, nxt);
        lput(cur);
        plane_xy(pn);
        if (!a.az) {
            store(zz, pn);
        } else if (zz - 1 >= zs && zz - 1 < ze) {
            /* z pass of plane zz - 1 from planes zz - 2 (if any), zz - 1, zz */
            float o[4];
#pragma unroll
            for (int cI = 0; cI < 4; ++cI) o[cI] = inv_outer_at(pm[cI], pc[cI], pn[cI], zz - 1, Z);
            store(zz - 1, o);
        }
#pragma unroll
        for (int cI = 0; cI < 4; ++cI) {
            pm[cI] = pc[cI];
            pc[cI] = pn[cI];
        }
#pragma unroll
        for (int i = 0; i < NLD; ++i) cur[i] = nxt[i];
        __syncthreads();   /* s_y reads done before the next plane rewrites it */
    }
    /* the last plane of the volume has no successor */
    if (a.az && zl == Z - 1 && Z - 1 >= zs && Z - 1 < ze) {
        float o[4];
#pragma unroll
        for (int cI = 0; cI < 4; ++cI) o[cI] = inv_outer_at(pm[cI], pc[cI], 0.0f, Z - 1, Z);
        store(Z - 1, o);
    }
}

struct InvFused {
    using Args = InvFusedArgs;
    static constexpr int THREADS = 256;
    static const char* name() { return "inv_bspline_fused"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { inv_fused_body(a, t); }
};

/* ---------------- local intensity scale ---------------- */
__device__ __forceinline__ int64_t refl(int64_t i, int64_t n) {
    /* reflection about the edge sample (paddPixels, PixelData.hpp:1070-1176) */
    if (i < 0) return -i;
    if (i >= n) return 2 * (n - 1) - i;
    return i;
}

struct PadGeom {
    int64_t yd, xd, zd;       /* unpadded (downsampled) dims */
    int64_t py, px, pz;       /* pads                         */
    int64_t yp, xp, zp;       /* padded dims                  */
    int64_t yq;               /* row pitch (>= yp): the LIS array rounds it up to 32 floats so rows start on
                                 128-byte lines and tile stores cover whole lines; unpadded arrays use yp */
};

/* y running-sum mean with offset O over padded lines (LocalIntensityScale.hpp:311-363).
   Source sample: mode 0 = padded copy of the intensity, mode 1 = |A - padded intensity|
   (calc_abs_diff, LocalIntensityScale.hpp:188-196).  O == 0 means "axis inactive": copy only.  Output written to A (may alias the source).
   One lane per line (the recursion is sequential in y); 64-line x 32-sample tiles are
   staged through LDS so global loads and stores stay coalesced along y.  The per-line
   source offsets are computed once per block. */
/* x / D for a small odd D (1..13): q0 = RN(x * RN(1/D)), the residual x - q0 * D exactly by
   FMA, one correction step, and the sign of x for zero results.  Equal to the IEEE quotient
   x / (float)D for every finite float x (checked over all 2^32 bit patterns for D = 1, 3, ..., 13). */
template <int D>
__device__ __forceinline__ float div_odd(float x) {
    constexpr float r = 1.0f / (float)D;
    const float q0 = x * r;
    const float e = fmaf(-q0, (float)D, x);
    return copysignf(fmaf(e, r, q0), x);
}

constexpr int MEANY_LINES = 64;
template <int O>
__device__ __forceinline__ void mean_y_body(int mode, const float* __restrict__ inten, float* A, const PadGeom& g, int64_t blk) {
    constexpr int TY = 32;          /* samples per tile                     */
    constexpr int NL = MEANY_LINES; /* lines per block, one per thread      */
    constexpr int H = 2 * O + 2;
    constexpr int P = TY + H + 1;   /* odd row pitch: lane-per-row accesses are conflict free */
    __shared__ float tile[NL][P];
    __shared__ int64_t src_row[NL];   /* inten offset of the (reflected) source row of each line */
    const int lane = threadIdx.x;
    const int64_t n_lines = g.xp * g.zp;
    const int64_t line0 = blk * NL;
    const int64_t me = line0 + lane;
    const bool valid = me < n_lines;
    const int yn = (int)g.yp, ydn = (int)g.yd, py = (int)g.py;
    const int64_t yq = g.yq;
    const float d = (float)(2 * O + 1);
    if (valid) {
        const int64_t xp = me % g.xp, zp = me / g.xp;
        src_row[lane] = (refl(zp - g.pz, g.zd) * g.xd + refl(xp - g.px, g.xd)) * g.yd;
    }
    __syncthreads();
    const int nl = (int)min((int64_t)NL, n_lines - line0);
    auto src = [&](int l, int yy) -> float {
        int ys = yy - py;
        if (ys < 0) ys = -ys;
        if (ys >= ydn) ys = 2 * (ydn - 1) - ys;
        const float b = inten[src_row[l] + ys];
        if (mode == 0) return b;
        return fabsf(A[(line0 + l) * yq + yy] - b);
    };
    float h[H > 0 ? H : 1];
    for (int i = 0; i < H; ++i) h[i] = 0.0f;
    float t = 0.0f;
    /* pre-consume samples [0, O) */
    if (valid) {
        for (int j = 0; j < O; ++j) {
            t += src(lane, j);
            for (int i = 0; i + 1 < H; ++i) h[i] = h[i + 1];
            h[H - 1] = t;
        }
    }
    /* Interior tiles (every output q has q > O and q + O < yn, every sample lies inside the
       unpadded line) run in two phases.  A: each lane turns its line's TY samples into running
       sums S (the only sequential dependence), stored after the H previous sums.  B: output
       q = p0 + k is (-S(q-O-1) + S(q+O)) / d in the reference's order, all independent, computed
       by lanes laid out along y so the stores are coalesced straight from registers. */
    const int64_t int_lo = max(32, py - O), int_hi = min((int64_t)yn, (int64_t)py + ydn) - O - TY;
    const int hb = lane >> 5, kk = lane & 31;
    for (int p0 = 0; p0 < yn; p0 += TY) {
        if (O > 0 && p0 >= int_lo && p0 <= int_hi) {
            /* loads go out in batches of LB before the first is used (lines past the end of the
               block's range read its last line and are not stored) */
            constexpr int LB = 32;
            const int64_t ys0 = p0 + O - py + kk;
            const float* arow = A + line0 * yq + p0 + O + kk;
#pragma unroll 1
            for (int i0 = 0; i0 < NL / 2; i0 += LB) {
                float vv[LB];
#pragma unroll
                for (int i = 0; i < LB; ++i) vv[i] = inten[src_row[min(2 * (i0 + i) + hb, nl - 1)] + ys0];
                if (mode != 0) {
                    float aa[LB];
#pragma unroll
                    for (int i = 0; i < LB; ++i) aa[i] = arow[(int64_t)min(2 * (i0 + i) + hb, nl - 1) * yq];
#pragma unroll
                    for (int i = 0; i < LB; ++i) vv[i] = fabsf(aa[i] - vv[i]);
                }
#pragma unroll
                for (int i = 0; i < LB; ++i)
                    if (2 * (i0 + i) + hb < nl) tile[2 * (i0 + i) + hb][H + kk] = vv[i];
            }
            __syncthreads();
            if (valid) {
                float* row = tile[lane];
                float xs[TY];
#pragma unroll
                for (int k = 0; k < TY; ++k) xs[k] = row[H + k];
#pragma unroll
                for (int m = 0; m < H; ++m) row[m] = h[m];
#pragma unroll
                for (int k = 0; k < TY; ++k) {
                    t += xs[k];
                    row[H + k] = t;
#pragma unroll
                    for (int i = 0; i + 1 < H; ++i) h[i] = h[i + 1];
                    h[H - 1] = t;
                }
            }
            __syncthreads();
            float* orow = A + line0 * yq + p0 + kk;
#pragma unroll 8
            for (int i = 0; i < NL / 2; ++i) {
                const int l = 2 * i + hb;
                float v = div_odd<2 * O + 1>(-tile[l][kk + 1]);
                v += div_odd<2 * O + 1>(tile[l][H + kk]);
                if (l < nl) orow[(int64_t)l * yq] = v;
            }
            __syncthreads();
            continue;
        }
        /* inputs for outputs [p0, p0+TY) are samples [p0+O, p0+O+TY): NL lines x TY samples */
        for (int idx = lane; idx < NL * TY; idx += NL) {
            const int l = idx / TY, k = idx % TY, yy = p0 + O + k;
            if (l < nl && yy < yn) tile[l][k] = src(l, yy);
        }
        __syncthreads();
        if (valid) {
            for (int k = 0; k < TY; ++k) {
                const int q = p0 + k;
                if (q >= yn) break;
                if (O == 0) continue;   /* copy: the tile already holds the samples */
                const float x = tile[lane][k];
                if (q + O < yn) {
                    t += x;
                    for (int i = 0; i + 1 < H; ++i) h[i] = h[i + 1];
                    h[H - 1] = t;
                }
                if (q < yn - O) {
                    float v = (q <= O) ? 0.0f : -h[0] / d;
                    v += h[H - 1] / d;
                    if (q >= 1 && q <= O) v = (float)((double)v * ((double)d / ((double)(q + O) + 1.0)));
                    if (q == 0) v = (float)((double)v * ((double)d / ((double)O + 1.0)));
                    tile[lane][k] = v;
                }
            }
        }
        __syncthreads();
        for (int idx = lane; idx < NL * TY; idx += NL) {
            const int l = idx / TY, k = idx % TY, q = p0 + k;
            if (l < nl && q < yn - O) A[(line0 + l) * yq + q] = tile[l][k];
        }
        __syncthreads();
    }
    /* right-hand end (LocalIntensityScale.hpp:346-352) */
    if (valid && O > 0) {
        const float tot = h[H - 1];
        for (int r = 0; r < O; ++r) {
            const int q = yn - O + r;
            if (q < 0) continue;
            float v = (q <= O) ? 0.0f : -h[1 + r] / d;
            const float cnt = (float)(r + 1);
            v *= d;
            v += tot;
            v = (float)((double)v * (1.0 / (double)(d - cnt)));
            if (q >= 1 && q <= O) v = (float)((double)v * ((double)d / ((double)(q + O) + 1.0)));
            if (q == 0) v = (float)((double)v * ((double)d / ((double)O + 1.0)));
            A[me * yq + q] = v;
        }
    }
}

/* x / z running-sum mean with offset O (LocalIntensityScale.hpp:365-494), in place;
   thread = one y of one line along the outer axis (length n, stride s) */
template <int O>
__device__ __forceinline__ void mean_outer_body(float* A, int64_t y_num, int64_t n, int64_t s, int64_t outer_stride, int64_t bx,
                                                int64_t by) {
    constexpr int H = 2 * O + 2;
    constexpr int W = 2 * O + 1;
    const int64_t y = bx * blockDim.x + threadIdx.x;
    if (y >= y_num) return;
    float* base = A + by * outer_stride + y;
    float h[H];
    for (int i = 0; i < H; ++i) h[i] = 0.0f;
    float c = 0.0f;
    for (int64_t j = 0; j < n; ++j) {
        const float x = base[j * s];
        c = (j == 0) ? x : x + c;
        for (int i = 0; i + 1 < H; ++i) h[i] = h[i + 1];
        h[H - 1] = c;
        const int64_t i = j - O;
        if (i < 0) continue;
        if (i <= O) {
            base[i * s] = c / (float)(i + O + 1);
        } else if (i < n - O) {
            base[i * s] = (c - h[0]) / (float)W;
        }
    }
    /* right-hand end: (C[last] - C[i-O-1]) / (n - i + O) */
    for (int r = 0; r < O; ++r) {
        const int64_t i = n - O + r;
        base[i * s] = (h[H - 1] - h[1 + r]) / (float)(n - i + O);
    }
}

/* the running mean of one pair of lines (two y, 8-byte accesses) along an outer axis of length
   n (stride sv in float2), D steps loaded ahead; emit(i, mean pair) for every output i.  Step j
   reads sample j and emits output j - O, so an emit that stores in place never overwrites a
   sample still to be read */
template <int O, int D, typename Emit>
__device__ __forceinline__ void mean_outer2_walk(const float2* base, int64_t n, int64_t sv, Emit emit) {
    constexpr int H = 2 * O + 2;
    constexpr int W = 2 * O + 1;
    float ha[H], hb[H];
#pragma unroll
    for (int i = 0; i < H; ++i) ha[i] = hb[i] = 0.0f;
    float ca = 0.0f, cb = 0.0f;
    for (int64_t j0 = 0; j0 < n; j0 += D) {
        float2 buf[D];
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (j0 + d < n) buf[d] = base[(j0 + d) * sv];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t j = j0 + d;
            if (j >= n) break;
            ca = (j == 0) ? buf[d].x : buf[d].x + ca;
            cb = (j == 0) ? buf[d].y : buf[d].y + cb;
#pragma unroll
            for (int i = 0; i + 1 < H; ++i) {
                ha[i] = ha[i + 1];
                hb[i] = hb[i + 1];
            }
            ha[H - 1] = ca;
            hb[H - 1] = cb;
            const int64_t i = j - O;
            if (i < 0) continue;
            if (i <= O) {
                const float dv = (float)(i + O + 1);
                emit(i, make_float2(ca / dv, cb / dv));
            } else if (i < n - O) {
                emit(i, make_float2(div_odd<W>(ca - ha[0]), div_odd<W>(cb - hb[0])));
            }
        }
    }
#pragma unroll
    for (int r = 0; r < O; ++r) {
        const int64_t i = n - O + r;
        const float dv = (float)(n - i + O);
        emit(i, make_float2((ha[H - 1] - ha[1 + r]) / dv, (hb[H - 1] - hb[1 + r]) / dv));
    }
}

/* the pass in place over lines of pairs numbered densely over (line, y pair), so rows whose
   length is not a multiple of the block still fill the blocks; lane counts stay below 2^32 (the
   launcher checks) */
template <int O, int D>
__device__ __forceinline__ void mean_outer2_body(float* A, int64_t y_num, int64_t n, int64_t s, int64_t outer_stride, int64_t lanes,
                                                 int64_t t) {
    const int64_t lane = t * blockDim.x + threadIdx.x;
    if (lane >= lanes) return;
    const uint32_t ny2 = (uint32_t)(y_num / 2);
    const uint32_t by = (uint32_t)lane / ny2, y2 = (uint32_t)lane - by * ny2;
    float2* base = (float2*)(A + by * outer_stride + 2 * y2);
    const int64_t sv = s / 2;
    mean_outer2_walk<O, D>(base, n, sv, [&](int64_t i, float2 v) { base[i * sv] = v; });
}

template <int O>
struct MeanY {
    struct Args {
        int mode;
        const float* inten;
        float* A;
        PadGeom g;
    };
    static constexpr int THREADS = MEANY_LINES;
    static const char* name() { return "lis_mean_y"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { mean_y_body<O>(a.mode, a.inten, a.A, a.g, t); }
};

struct OuterArgs {
    float* A;
    int64_t y_num, n, s, outer_stride, gx;
};
template <int O>
struct MeanOuter {
    using Args = OuterArgs;
    static constexpr int THREADS = 256;
    static const char* name() { return "lis_mean_outer"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { mean_outer_body<O>(a.A, a.y_num, a.n, a.s, a.outer_stride, t % a.gx, t / a.gx); }
};
template <int O, int D>
struct MeanOuter2 {
    using Args = OuterArgs;
    static constexpr int THREADS = 256;
    static const char* name() { return "lis_mean_outer2"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        mean_outer2_body<O, D>(a.A, a.y_num, a.n, a.s, a.outer_stride, a.gx, t);
    }
};

template <int O>
void mean_y(aprh_context* ctx, int mode, const float* inten, float* A, const PadGeom& g) {
    const int64_t lines = g.xp * g.zp;
    launch_tiles<MeanY<O>>(ctx, {mode, inten, A, g}, cdiv(lines, MEANY_LINES));
}

void mean_y_dispatch(aprh_context* ctx, int O, int mode, const float* inten, float* A, const PadGeom& g) {
    switch (O) {
        case 0: mean_y<0>(ctx, mode, inten, A, g); break;
        case 1: mean_y<1>(ctx, mode, inten, A, g); break;
        case 2: mean_y<2>(ctx, mode, inten, A, g); break;
        case 3: mean_y<3>(ctx, mode, inten, A, g); break;
        case 4: mean_y<4>(ctx, mode, inten, A, g); break;
        case 5: mean_y<5>(ctx, mode, inten, A, g); break;
        case 6: mean_y<6>(ctx, mode, inten, A, g); break;
        default: throw Error(APRH_ERR_UNSUPPORTED, "mean window offset > 6");
    }
}

template <int O>
void mean_outer(aprh_context* ctx, float* A, int64_t y_num, int64_t n, int64_t s, int64_t outer_stride, int64_t n_outer) {
    const int64_t lanes2 = y_num / 2 * n_outer;
    if (y_num % 2 == 0 && s % 2 == 0 && outer_stride % 2 == 0 && ((uintptr_t)A & 7u) == 0 && lanes2 < (int64_t(1) << 32)) {
        /* the gx field carries the lane count of the dense numbering; a batch plan also times
           4 and 16 steps of look-ahead */
        const OuterArgs oa{A, y_num, n, s, outer_stride, lanes2};
        const int64_t tiles = cdiv(lanes2, 256);
        if (ctx->recorder)
            launch_tiles_alts(ctx,
                              {tile_alt<MeanOuter2<O, 4>>(oa, tiles, "look-ahead 4"), tile_alt<MeanOuter2<O, 8>>(oa, tiles, "look-ahead 8"),
                               tile_alt<MeanOuter2<O, 16>>(oa, tiles, "look-ahead 16")},
                              1);
        else
            launch_tiles<MeanOuter2<O, 8>>(ctx, oa, tiles);
    } else {
        const int64_t gx = cdiv(y_num, 256);
        launch_tiles<MeanOuter<O>>(ctx, {A, y_num, n, s, outer_stride, gx}, gx * n_outer);
    }
}

void mean_outer_dispatch(aprh_context* ctx, int O, float* A, int64_t y_num, int64_t n, int64_t s, int64_t outer_stride,
                         int64_t n_outer) {
    if (n < 2 * O + 2) throw Error(APRH_ERR_UNSUPPORTED, "mean window wider than the axis");
    switch (O) {
        case 1: mean_outer<1>(ctx, A, y_num, n, s, outer_stride, n_outer); break;
        case 2: mean_outer<2>(ctx, A, y_num, n, s, outer_stride, n_outer); break;
        case 3: mean_outer<3>(ctx, A, y_num, n, s, outer_stride, n_outer); break;
        case 4: mean_outer<4>(ctx, A, y_num, n, s, outer_stride, n_outer); break;
        case 5: mean_outer<5>(ctx, A, y_num, n, s, outer_stride, n_outer); break;
        case 6: mean_outer<6>(ctx, A, y_num, n, s, outer_stride, n_outer); break;
        default: throw Error(APRH_ERR_UNSUPPORTED, "mean window offset must be 1..6");
    }
}

/* x mean over A (dims y,x,z): outer = z; z mean: outer = x */
void mean_x(aprh_context* ctx, int O, float* A, int64_t y, int64_t x, int64_t z) {
    mean_outer_dispatch(ctx, O, A, y, x, y, x * y, z);
}
void mean_z(aprh_context* ctx, int O, float* A, int64_t y, int64_t x, int64_t z) {
    mean_outer_dispatch(ctx, O, A, y, z, x * y, y, x);
}

/* applyParameters + computeLevels, one lane per downsampled cell */
template <typename G>
struct LevelsArgs {
    const float* A;
    PadGeom g;
    float var_rescale;
    const float* inten;
    G* grad;
    float* sigma_raw;
    float* sigma_out;
    uint8_t* level;
    const float* d_off;
    float Ip_th, sigma_th, sigma_th_max, grad_th, mult;
    G* grad_keep;
    int64_t Gs;   /* virtual grid */
    FastDiv f_yd, f_xd;
    int i32;
};
/* applyParameters and computeLevels of downsampled cell i from its rescaled scale s
   (threshold_gradient ComputeGradient.hpp:159-171, APRConverter.hpp:249-297,
   LocalParticleCellSet.hpp:128-171); ipth = Ip_th + bspline offset */
template <typename G>
__device__ __forceinline__ void levels_at(const LevelsArgs<G>& a, size_t i, float s, float ipth) {
    if (a.sigma_raw) a.sigma_raw[i] = s;
    G gv = a.grad[i];
    if (a.inten[i] <= ipth) gv = (G)0;
    if (s < a.sigma_th) s = (s < a.sigma_th_max) ? 60000.0f : a.sigma_th;
    if ((float)gv < a.grad_th) gv = (G)0;
    if (a.sigma_out) a.sigma_out[i] = s;
    if (a.grad_keep) a.grad_keep[i] = gv;
    const float L = (float)gv / s;
    const uint32_t u = f2u32(L * a.mult);
    a.level[i] = (uint8_t)(u == 0u ? 0 : 31 - __clz(u));
}

template <typename G>
__device__ __forceinline__ void levels_body(const LevelsArgs<G>& a, int64_t t) {
    const float* __restrict__ A = a.A;
    const PadGeom& g = a.g;
    const float var_rescale = a.var_rescale;
    const float ipth = a.Ip_th + *a.d_off;
    const size_t n = (size_t)(g.yd * g.xd * g.zd);
    for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < n; i += (size_t)a.Gs * blockDim.x) {
        int64_t y, x, z;
        if (a.i32) {
            const uint32_t r = a.f_yd.div((uint32_t)i), zz = a.f_xd.div(r);
            y = (int64_t)((uint32_t)i - r * (uint32_t)g.yd);
            x = (int64_t)(r - zz * (uint32_t)g.xd);
            z = zz;
        } else {
            y = (int64_t)(i % (size_t)g.yd);
            const int64_t r = (int64_t)(i / (size_t)g.yd);
            x = r % g.xd;
            z = r / g.xd;
        }
        /* rescale_var + unpad (LocalIntensityScale.hpp:173-184, PixelData.hpp:1189-1210) */
        levels_at(a, i, A[((z + g.pz) * g.xp + (x + g.px)) * g.yq + (y + g.py)] * var_rescale, ipth);
    }
}

/* the local intensity scale's last pass (the z mean of offset O, LocalIntensityScale.hpp:434-494)
   with the levels epilogue: lanes (x column, y pair) of the unpadded cells walk the padded z
   line, and each unpadded output, rescaled, goes to levels_at instead of back to A.  Lanes are
   numbered densely over (x, pair); pairs start at the even y at or below the first unpadded y,
   and the half of a pair outside the unpadded range is not used. */
template <int O, int D, typename G>
__device__ __forceinline__ void mean_z_levels_body(const LevelsArgs<G>& a, int64_t p0, int64_t npair, int64_t lanes, int64_t t) {
    const int64_t lane = t * blockDim.x + threadIdx.x;
    if (lane >= lanes) return;
    const PadGeom& g = a.g;
    const uint32_t xi = (uint32_t)lane / (uint32_t)npair, pi = (uint32_t)lane - xi * (uint32_t)npair;
    const int64_t yp0 = 2 * (p0 + pi);              /* padded y of the pair's first cell */
    const float2* base = (const float2*)(a.A + (xi + g.px) * g.yq + yp0);
    const int64_t sv = g.xp * g.yq / 2;
    const float ipth = a.Ip_th + *a.d_off;
    const float vr = a.var_rescale;
    const int64_t y0 = yp0 - g.py;                  /* unpadded y of the first cell (may be -1) */
    const bool in0 = y0 >= 0 && y0 < g.yd, in1 = y0 + 1 >= 0 && y0 + 1 < g.yd;
    mean_outer2_walk<O, D>(base, g.zp, sv, [&](int64_t i, float2 v) {
        const int64_t z = i - g.pz;
        if (z < 0 || z >= g.zd) return;
        const size_t row = (size_t)((z * g.xd + xi) * g.yd);
        if (in0) levels_at(a, row + (size_t)y0, v.x * vr, ipth);
        if (in1) levels_at(a, row + (size_t)(y0 + 1), v.y * vr, ipth);
    });
}

/* rescale_var + unpaddPixels (LocalIntensityScale.hpp:173-184, PixelData.hpp:1189-1210) */
__device__ __forceinline__ void unpad_rescale_body(const float* __restrict__ A, const PadGeom& g, float var_rescale, float* out,
                                                   int64_t G, int64_t t) {
    const size_t n = (size_t)(g.yd * g.xd * g.zd);
    for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < n; i += (size_t)G * blockDim.x) {
        const int64_t y = (int64_t)(i % (size_t)g.yd);
        const int64_t r = (int64_t)(i / (size_t)g.yd);
        const int64_t x = r % g.xd, z = r / g.xd;
        out[i] = A[((z + g.pz) * g.xp + (x + g.px)) * g.yq + (y + g.py)] * var_rescale;
    }
}

/* constant local intensity scale (LocalIntensityScale.hpp:131-148): a sequential double
   sum in index order, kept sequential (single lane) to reproduce its rounding */
__device__ __forceinline__ void constant_scale_sum_body(const float* __restrict__ m, size_t n, float* out_val) {
    if (threadIdx.x != 0) return;
    float mn = 660000.0f;
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const float t = m[i];
        sum += (double)t;
        if (t < mn) mn = t;
    }
    *out_val = (float)(sum / (double)(float)n - (double)mn);
}

__device__ __forceinline__ void fill_body(float* m, size_t n, const float* __restrict__ val, int64_t G, int64_t t) {
    const float v = *val;
    for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < n; i += (size_t)G * blockDim.x) m[i] = v;
}

/* tile bodies of the gradient, inverse B-spline and level stages */
struct GradArgs {
    const void* img;
    void* grad;
    float* mean;
    int64_t y_num, x_num, z_num, yd_num;
    float h2x, h2y, h2z;
    int accum;
    int64_t gx, xd_num, zd_num;
};
template <typename T>
struct GradDs {
    using Args = GradArgs;
    static constexpr int THREADS = 256;
    static const char* name() { return "gradient_ds"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        const int64_t bx = t % a.gx, r = t / a.gx;
        grad_ds_body<T>((const T*)a.img, (T*)a.grad, a.mean, a.y_num, a.x_num, a.z_num, a.yd_num, a.h2x, a.h2y, a.h2z, a.accum, bx,
                        r % a.xd_num, r / a.xd_num, a.xd_num);
    }
};
template <bool UNIT>
struct GradDsVec {
    using Args = GradArgs;
    static constexpr int THREADS = 256;
    static const char* name() { return "gradient_ds_vec"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        /* lanes numbered densely over (y group, x column pair, run of GZC z planes), so short y
           rows still fill the blocks */
        /* valid only for y_num % 8 == 0, which launch_grad_ds_u16 checks before choosing it */
        const int64_t ny = a.y_num / 8, nz = (a.zd_num + GZC - 1) / GZC;
        const int64_t L = t * blockDim.x + threadIdx.x;
        if (L >= ny * a.xd_num * nz) return;
        const int64_t yg = L % ny, r = L / ny;
        grad_ds_vec_body<UNIT>((const uint16_t*)a.img, (uint16_t*)a.grad, a.mean, a.y_num, a.x_num, a.z_num, a.yd_num, a.h2x, a.h2y,
                               a.h2z, a.accum, yg, r % a.xd_num, r / a.xd_num, a.xd_num, a.zd_num);
    }
};
template <typename G>
struct MeanZLevelsArgs {
    LevelsArgs<G> la;
    int64_t p0, npair, lanes;
};
template <int O, typename G, int D = 8>
struct MeanZLevels {
    using Args = MeanZLevelsArgs<G>;
    static constexpr int THREADS = 256;
    static const char* name() { return "lis_mean_z_levels"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { mean_z_levels_body<O, D, G>(a.la, a.p0, a.npair, a.lanes, t); }
};
template <typename G>
struct Levels {
    using Args = LevelsArgs<G>;
    static constexpr int THREADS = 256;
    static const char* name() { return "levels"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { levels_body<G>(a, t); }
};
struct UnpadRescale {
    struct Args {
        const float* A;
        PadGeom g;
        float var_rescale;
        float* out;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "unpad_rescale"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { unpad_rescale_body(a.A, a.g, a.var_rescale, a.out, a.G, t); }
};
struct ConstantScaleSum {
    struct Args {
        const float* m;
        size_t n;
        float* out_val;
    };
    static constexpr int THREADS = 64;
    static const char* name() { return "constant_scale_sum"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t) { constant_scale_sum_body(a.m, a.n, a.out_val); }
};
struct Fill {
    struct Args {
        float* m;
        size_t n;
        const float* val;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "fill"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { fill_body(a.m, a.n, a.val, a.G, t); }
};

PadGeom pad_geom(int64_t yd, int64_t xd, int64_t zd, const LisParams& lp) {
    PadGeom g;
    g.yd = yd; g.xd = xd; g.zd = zd;
    g.py = lp.pad[0]; g.px = lp.pad[1]; g.pz = lp.pad[2];
    g.yp = yd + 2 * g.py; g.xp = xd + 2 * g.px; g.zp = zd + 2 * g.pz;
    const int64_t a = std::max(1, lp.pitch_align);
    g.yq = (g.yp + a - 1) / a * a;
    return g;
}

}  // namespace

void launch_grad_ds_u16(aprh_context* ctx, const uint16_t* img, uint16_t* grad, float* mean, int64_t y, int64_t x, int64_t z,
                        float hx, float hy, float hz, bool accum) {
    const int64_t yd = cdiv(y, 2), xd = cdiv(x, 2), zd = cdiv(z, 2);
    const bool aligned = (((uintptr_t)img | (uintptr_t)grad | (uintptr_t)mean) & 15u) == 0;
    if (y % 8 == 0 && aligned) {
        /* GradDsVec owns 8 consecutive y per lane and numbers its lanes densely, so it needs
           y % 8 == 0 (tested here) and takes no row-tile count (gx = 0) */
        const bool unit = hx == 1.0f && hy == 1.0f && hz == 1.0f;
        GradArgs a{img, grad, mean, y, x, z, yd, 2.0f * hx, 2.0f * hy, 2.0f * hz, accum ? 1 : 0, 0, xd, zd};
        const int64_t tiles = cdiv((y / 8) * xd * cdiv(zd, GZC), 256);   /* 256 lanes per tile */
        if (unit) launch_tiles<GradDsVec<true>>(ctx, a, tiles);
        else launch_tiles<GradDsVec<false>>(ctx, a, tiles);
    } else {
        const int64_t gx = cdiv(yd, 256);
        launch_tiles<GradDs<uint16_t>>(ctx, GradArgs{img, grad, mean, y, x, z, yd, 2.0f * hx, 2.0f * hy, 2.0f * hz, accum ? 1 : 0, gx, xd, zd},
                                       gx * xd * zd);
    }
}

int inv_bspline_launches(bool ay, bool ax, bool az) { return (ay || ax || az) ? 1 : 0; }

float* launch_inv_bspline_pp(aprh_context* ctx, float* a, float* b, int64_t y, int64_t x, int64_t z, bool ay, bool ax, bool az) {
    if (ay || ax || az) {
        /* one sweep for all active axes: a -> b */
        if ((ay && y < 2) || (ax && x < 2) || (az && z < 2)) throw Error(APRH_ERR_UNSUPPORTED, "inverse B-spline over an axis shorter than 2");
        const int64_t nty = cdiv(y, INV_TY), ntx = cdiv(x, INV_TX), ntz = cdiv(z, INV_ZC);
        launch_tiles<InvFused>(ctx, InvFusedArgs{a, b, y, x, z, ay ? 1 : 0, ax ? 1 : 0, az ? 1 : 0, nty, ntx}, nty * ntx * ntz);
        return b;
    }
    return a;
}

void launch_inv_bspline(aprh_context* ctx, float* m, float* tmp, int64_t y, int64_t x, int64_t z, bool ay, bool ax, bool az) {
    /* m -> tmp, copied back so the result ends in m */
    float* fin = launch_inv_bspline_pp(ctx, m, tmp, y, x, z, ay, ax, az);
    if (fin != m) launch_copy(ctx, m, fin, (size_t)(y * x * z) * 4);
}

LisParams lis_params(const aprh_params& par, int64_t yd, int64_t xd, int64_t zd) {
    /* get_window_alt (LocalIntensityScale.hpp:250-303) */
    static const double r3[6] = {12.8214, 26.1256, 40.2795, 23.3692, 36.2061, 27.0385};
    static const double r2[6] = {13.2421, 28.7069, 52.0385, 24.4272, 34.9565, 21.1891};
    static const double r1[6] = {13.9040, 31.2843, 57.4037, 30.3767, 45.9930, 29.0890};
    static const int w1[] = {1, 1, 1, 2, 2, 3}, w2[] = {2, 3, 4, 4, 5, 6};
    int psf_ind = (int)std::max((float)(std::round((double)(par.psfx / par.dx)) - 1), 0.0f);
    psf_ind = std::min(psf_ind, 5);
    const int wv = std::max(w1[psf_ind], w2[psf_ind]);
    LisParams lp;
    const int64_t dims[3] = {yd, xd, zd};
    int nact = 0;
    for (int a = 0; a < 3; ++a) {
        lp.active[a] = dims[a] > wv;
        lp.win[a] = lp.active[a] ? w1[psf_ind] : 0;
        lp.win[a + 3] = lp.active[a] ? w2[psf_ind] : 0;
        nact += lp.active[a];
        /* paddPixels clamps the pad to dim-1 (PixelData.hpp:1083-1099) */
        int p = std::max(lp.win[a], lp.win[a + 3]);
        lp.pad[a] = dims[a] > 1 ? (int)std::min<int64_t>(p, dims[a] - 1) : 0;
    }
    lp.var_rescale = (float)(nact == 3 ? r3[psf_ind] : nact == 2 ? r2[psf_ind] : r1[psf_ind]);
    lp.pitch_align = 32;   /* 128-byte rows for the padded array launch_lis fills */
    return lp;
}

void launch_sat_mean(aprh_context* ctx, float* m, int64_t y, int64_t x, int64_t z, int axis, int off) {
    if (axis == 0) {
        LisParams lp{};
        PadGeom g = pad_geom(y, x, z, lp);
        (void)g;
        /* y mean of an unpadded array: treat it as a padded array with zero pads whose source is itself */
        PadGeom gg;
        gg.yd = y; gg.xd = x; gg.zd = z; gg.py = gg.px = gg.pz = 0; gg.yp = y; gg.xp = x; gg.zp = z; gg.yq = y;
        if (y < 2 * off + 2) throw Error(APRH_ERR_UNSUPPORTED, "mean window wider than the axis");
        mean_y_dispatch(ctx, off, 0, m, m, gg);
    } else if (axis == 1) {
        mean_x(ctx, off, m, y, x, z);
    } else {
        mean_z(ctx, off, m, y, x, z);
    }
}

void launch_lis(aprh_context* ctx, const float* intensity, float* sigma_padded, float* /*unused*/, int64_t yd, int64_t xd,
                int64_t zd, const LisParams& lp) {
    PadGeom g = pad_geom(yd, xd, zd, lp);
    float* A = sigma_padded;
    /* first mean (windows win[0..2]) of the padded intensity */
    mean_y_dispatch(ctx, lp.active[0] ? lp.win[0] : 0, 0, intensity, A, g);
    /* x / z passes over the yp valid columns of rows with pitch yq */
    auto mx = [&](int O) { mean_outer_dispatch(ctx, O, A, g.yp, g.xp, g.yq, g.xp * g.yq, g.zp); };
    auto mz = [&](int O) { mean_outer_dispatch(ctx, O, A, g.yp, g.zp, g.xp * g.yq, g.yq, g.xp); };
    if (lp.active[1]) mx(lp.win[1]);
    if (lp.active[2]) mz(lp.win[2]);
    /* |mean - intensity| then the second mean (windows win[3..5]) */
    mean_y_dispatch(ctx, lp.active[0] ? lp.win[3] : 0, 1, intensity, A, g);
    if (lp.active[1]) mx(lp.win[4]);
    /* with fuse_levels the last z mean runs in the levels launch (launch_levels_padded) */
    if (lp.active[2] && !lp.fuse_levels) mz(lp.win[5]);
}

bool lis_can_fuse_levels(const LisParams& lp, int64_t yd, int64_t xd, int64_t zd) {
    if (!lp.active[2] || lp.win[5] < 1 || lp.win[5] > 6) return false;
    PadGeom g = pad_geom(yd, xd, zd, lp);
    if (g.yq % 2 != 0 || g.zp < 2 * lp.win[5] + 2) return false;
    const int64_t npair = (g.py + g.yd + 1) / 2 - g.py / 2;
    return npair * g.xd < (int64_t(1) << 32);
}

/* the fused last mean with 8 steps of look-ahead; a batch plan (u16) also times 4 and 16 */
template <int O, typename G>
void launch_mean_z_levels(aprh_context* ctx, const MeanZLevelsArgs<G>& za, int64_t tiles) {
    if constexpr (std::is_same<G, uint16_t>::value) {
        if (ctx->recorder) {
            launch_tiles_alts(ctx,
                              {tile_alt<MeanZLevels<O, G, 4>>(za, tiles, "look-ahead 4"), tile_alt<MeanZLevels<O, G, 8>>(za, tiles, "look-ahead 8"),
                               tile_alt<MeanZLevels<O, G, 16>>(za, tiles, "look-ahead 16")},
                              1);
            return;
        }
    }
    launch_tiles<MeanZLevels<O, G>>(ctx, za, tiles);
}

template <typename G>
void launch_levels_padded(aprh_context* ctx, const float* A, const LisParams& lp, int64_t yd, int64_t xd, int64_t zd,
                          const float* inten, G* grad, float* sigma_raw, float* sigma_out, uint8_t* level,
                          const aprh_params& par, const float* d_offset, float mult, G* grad_keep) {
    PadGeom g = pad_geom(yd, xd, zd, lp);
    const int64_t Gs = stride_tiles(yd * xd * zd, 256 * 4);
    const int i32 = fits32(yd * xd * zd) ? 1 : 0;
    const LevelsArgs<G> la{A, g, lp.var_rescale, inten, grad, sigma_raw, sigma_out, level, d_offset, par.Ip_th,
                           par.sigma_th, par.sigma_th_max, par.grad_th, mult, grad_keep, Gs,
                           i32 ? FastDiv((uint64_t)yd) : FastDiv(), i32 ? FastDiv((uint64_t)xd) : FastDiv(), i32};
    if (lp.fuse_levels) {
        /* the z mean of the second window (left out by launch_lis) with the levels epilogue */
        const int64_t p0 = g.py / 2, npair = (g.py + g.yd + 1) / 2 - p0, lanes = npair * g.xd;
        const MeanZLevelsArgs<G> za{la, p0, npair, lanes};
        const int64_t tiles = cdiv(lanes, 256);
        switch (lp.win[5]) {
            case 1: launch_mean_z_levels<1, G>(ctx, za, tiles); break;
            case 2: launch_mean_z_levels<2, G>(ctx, za, tiles); break;
            case 3: launch_mean_z_levels<3, G>(ctx, za, tiles); break;
            case 4: launch_mean_z_levels<4, G>(ctx, za, tiles); break;
            case 5: launch_mean_z_levels<5, G>(ctx, za, tiles); break;
            case 6: launch_mean_z_levels<6, G>(ctx, za, tiles); break;
            default: throw Error(APRH_ERR_UNSUPPORTED, "mean window offset must be 1..6");
        }
        return;
    }
    launch_tiles<Levels<G>>(ctx, la, Gs);
}

template void launch_levels_padded<uint16_t>(aprh_context*, const float*, const LisParams&, int64_t, int64_t, int64_t, const float*,
                                             uint16_t*, float*, float*, uint8_t*, const aprh_params&, const float*, float, uint16_t*);
template void launch_levels_padded<float>(aprh_context*, const float*, const LisParams&, int64_t, int64_t, int64_t, const float*, float*,
                                          float*, float*, uint8_t*, const aprh_params&, const float*, float, float*);

void launch_grad_ds_f32(aprh_context* ctx, const float* img, float* grad, float* mean, int64_t y, int64_t x, int64_t z, float hx,
                        float hy, float hz, bool accum) {
    const int64_t yd = cdiv(y, 2), xd = cdiv(x, 2), zd = cdiv(z, 2);
    const int64_t gx = cdiv(yd, 256);
    launch_tiles<GradDs<float>>(ctx, GradArgs{img, grad, mean, y, x, z, yd, 2.0f * hx, 2.0f * hy, 2.0f * hz, accum ? 1 : 0, gx, xd, zd},
                                gx * xd * zd);
}

void launch_unpad_rescale(aprh_context* ctx, const float* A, const LisParams& lp, int64_t yd, int64_t xd, int64_t zd, float* out) {
    PadGeom g = pad_geom(yd, xd, zd, lp);
    const int64_t G = stride_tiles(yd * xd * zd, 256 * 4);
    launch_tiles<UnpadRescale>(ctx, {A, g, lp.var_rescale, out, G}, G);
}

namespace {
struct U8ToF32 {
    struct Args {
        const uint8_t* in;
        float* out;
        size_t n;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "u8_to_f32"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < a.n; i += (size_t)a.G * blockDim.x) a.out[i] = (float)a.in[i];
    }
};
}  // namespace

void launch_u8_to_f32(aprh_context* ctx, const uint8_t* in, float* out, size_t n) {
    const int64_t G = stride_tiles((int64_t)n, 256 * 4);
    launch_tiles<U8ToF32>(ctx, {in, out, n, G}, G);
}

void launch_constant_scale(aprh_context* ctx, float* m, size_t n) {
    /* a recorded (batched) volume needs its own value slot: the plan's pool keeps every block */
    Buf own;
    if (ctx->recorder) own.reset(ctx->pool, 256);
    float* d_val = ctx->recorder ? own.as<float>() : (float*)ctx->d_flags + 32;
    launch_tiles<ConstantScaleSum>(ctx, {m, n, d_val}, 1);
    const int64_t G = stride_tiles((int64_t)n, 256 * 4);
    launch_tiles<Fill>(ctx, {m, n, d_val, G}, G);
}

size_t lis_padded_elems(int64_t yd, int64_t xd, int64_t zd, const LisParams& lp) {
    return (size_t)(pad_geom(yd, xd, zd, lp).yq * (xd + 2 * lp.pad[1]) * (zd + 2 * lp.pad[2]));
}

}  // namespace aprh
