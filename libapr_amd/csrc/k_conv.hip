/*
 * k_conv.hip -- APR convolution over level tiles, with the tree means and the
 * Richardson-Lucy steps that surround it.
 *
 * APRFilter::convolve_pencil (APRFilter.hpp:1094-1216) evaluates, for every
 * particle of level l, the level's stencil over the isotropic neighbourhood at
 * level l: level-l particles, coarser particles upsampled, and tree means where
 * finer particles live (update_same_level / update_higher_level /
 * update_lower_level, :743-849; reflective or zero ends, :384-504), summing taps
 * in (z, x, y) order (run_convolution_pencil, :705-736).
 *
 * Here one launch per level walks tiles of TX x 64 cells (x, y) through a run of
 * z planes.  For each plane the block assembles the tile and its halo in LDS:
 *   1. the coarser level's image V(l-1) upsampled (every cell not covered by a
 *      level-l particle or tree node lies inside a coarser particle),
 *   2. the level's particles of the halo rows, found through a per-row index of
 *      64-cell y blocks (particles with y in a block are contiguous),
 *   3. the level's tree means the same way,
 *   4. y ends mirrored in LDS; x / z ends read the reflected source row / plane.
 * The assembled interior is also written to V(l), the next level's coarse
 * source (tree cells included: the finer level covers them with its own
 * particles or tree nodes).  Planes sit in an LDS ring of the stencil's z size;
 * once a plane's z neighbourhood is in LDS, one lane per particle of its rows
 * sums the taps:
 *   - reference order: every tap in (z, x, y) order, multiply then add;
 *   - separable (rank-1 stencils): a y pass and an x pass per plane into the
 *     ring, the z pass at the particles.
 * The result goes through an epilogue: store, or Richardson-Lucy's
 * input / blur (APRNumerics.hpp:553, 569), or u *= estimate with the partial
 * sums of the relative-update norm.
 *
 * Tree means (APRTreeNumerics::fill_tree_mean, APRTreeNumerics.hpp:63-202): the
 * top tree level's children are all particles of l_max, and the four particle
 * rows under a tree row hold exactly the children of its nodes, two per node
 * (one at an odd y end).  Its means are therefore read straight from the
 * particle rows, children in (z, x, y) order, after a one-time check of that row
 * shape per APR.
 */
#include <algorithm>
#include <cmath>

#include "aprh_internal.hpp"
#include "tiles.hpp"

namespace aprh {

namespace {

constexpr int TY = 64;        /* tile width along y: one wave         */
constexpr int CT = 256;       /* threads per block                    */
constexpr int NWAVE = CT / 64;

struct LevelDesc {
    int64_t Y, X, Z;              /* level dims                                           */
    int64_t row_begin;            /* xz_end index of row (z=0, x=0)                        */
    int64_t t_row_begin;          /* t_xz_end index of tree row (0, 0); -1: no tree nodes */
    int64_t cY, cX;               /* coarse image dims (level - 1)                         */
    int64_t yidx_off, t_yidx_off; /* offsets into the y-block indexes                      */
    int nyb;                      /* y blocks of the level: ceil(Y / 64)                   */
    int ntx, nty, ntz, zc;        /* tiles along x, y, z; planes per z tile                */
    int sy, sx, sz;               /* stencil size                                          */
    int64_t w_off;                /* weights: reference order sz*sx*sy; separable fy|fx|fz */
};

template <typename TIn>
struct ConvArgs {
    LevelDesc L;
    const uint64_t* xz_end;
    const uint16_t* y_vec;
    const TIn* in;
    const uint64_t* t_xz_end;
    const uint16_t* t_y_vec;
    const float* tree;
    const uint16_t* yidx;
    const uint16_t* t_yidx;
    const float* coarse;      /* V(l - 1), or nullptr at l_min   */
    float* vout;              /* V(l) for the next level, or nullptr */
    const float* W;
    float* out;
    const void* epi_in;       /* epilogue 1: the RL input particles */
    double* part;             /* epilogue 2: 2 doubles per block    */
    int reflect;
    int epi;                  /* 0 store, 1 out = epi_in / val, 2 out *= val (+ norms) */
    int epi_in_u16;
};

__device__ __forceinline__ int64_t refl_c(int64_t q, int64_t n) {
    /* reflection about the edge sample (no repeat), clamped for lines shorter than the halo */
    if (q < 0) q = -q;
    if (q >= n) q = 2 * (n - 1) - q;
    return q < 0 ? 0 : (q >= n ? n - 1 : q);
}

/* particles of row `rid` (an xz_end index): [b, e), and the part [s0, s1) with y in block yb */
__device__ __forceinline__ void row_block(const uint64_t* __restrict__ xz, const uint16_t* __restrict__ yidx, int64_t yrow, int nyb,
                                          int64_t rid, int yb, uint64_t& b, uint64_t& e, uint64_t& s0, uint64_t& s1) {
    if (nyb <= 1) {
        b = xz[rid - 1];
        e = xz[rid];
        s0 = b;
        s1 = e;
        return;
    }
    /* both index entries are read (clamped to the row's nyb - 1 entries) together with the row
       ends, so the three loads are in flight at once */
    const uint16_t* ix = yidx + yrow * (int64_t)(nyb - 1);
    const uint64_t m0 = yb > 0 ? ~(uint64_t)0 : 0, m1 = yb + 1 < nyb ? ~(uint64_t)0 : 0;
    const uint16_t v0 = ix[yb > 0 ? yb - 1 : 0], v1 = ix[yb < nyb - 2 ? yb : nyb - 2];
    b = xz[rid - 1];
    e = xz[rid];
    /* selections as masks, not branches, so no wait falls between the loads */
    s0 = b + ((uint64_t)v0 & m0);
    s1 = e + ((b + v1 - e) & m1);
}

template <typename T> __device__ __forceinline__ float tof(T v) { return (float)v; }

/* LDS carve-up of one block */
struct Lds {
    float* ring;      /* exact: sz planes of WX x WYS; separable: the B ring, sz planes of TX x 64 */
    float* V;         /* separable: the assembled plane (WX x WYS); exact: ring slot             */
    float* A;         /* separable: y pass, WX x 64                                              */
    uint8_t* ylist;   /* sz x TX x 64 particle y offsets of interior rows                        */
    int* cnt;         /* sz x TX                                                                 */
    uint64_t* base;   /* sz x TX first particle index                                            */
    uint64_t* mlo;    /* 2 WX: first index to load of each window row (particle rows, tree rows) */
    int* mpre;        /* 2 WX + 1: exclusive prefix of the rows' load counts                     */
};

/* weights through the constant address space: uniform indices become scalar loads */
typedef const __attribute__((address_space(4))) float* CWeights;

/* sum over one window row of the reference-order taps: SY > 0 unrolled, 0 runtime sy */
template <int SY>
__device__ __forceinline__ float row_taps(const float* rw, CWeights w, int sy, float val) {
    if (SY > 0) {
#pragma unroll
        for (int iy = 0; iy < SY; ++iy) val += rw[iy] * w[iy];
    } else {
        for (int iy = 0; iy < sy; ++iy) val += rw[iy] * w[iy];
    }
    return val;
}

template <typename TIn, int TX, bool SEP, int SY>
__global__ __launch_bounds__(CT) void k_conv_tile(ConvArgs<TIn> a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const LevelDesc& L = a.L;
    const int hy = (L.sy - 1) / 2, hx = (L.sx - 1) / 2, hz = (L.sz - 1) / 2;
    const int WX = TX + 2 * hx, WY = TY + 2 * hy, WYS = WY;
    const int sz = L.sz;

    /* carve LDS (sizes mirror conv_lds_bytes); the 8-byte entries first */
    Lds s;
    uint8_t* p = smem;
    s.base = (uint64_t*)p;
    p += (size_t)sz * TX * 8;
    s.mlo = (uint64_t*)p;
    p += (size_t)2 * WX * 8;
    s.mpre = (int*)p;
    p += (size_t)(2 * WX + 2) * 4;
    s.cnt = (int*)p;
    p += (size_t)sz * TX * 4;
    s.ylist = p;
    p += (size_t)sz * TX * TY;
    if (SEP) {
        s.ring = (float*)p;
        p += (size_t)sz * TX * TY * 4;
        s.V = (float*)p;
        p += (size_t)WX * WYS * 4;
        s.A = (float*)p;
        p += (size_t)WX * TY * 4;
    } else {
        s.ring = (float*)p;
        p += (size_t)sz * WX * WYS * 4;
        s.V = nullptr;
        s.A = nullptr;
    }

    const int64_t ntiles = (int64_t)L.ntx * L.nty * L.ntz;
    const int64_t tile = xcd_block(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const int tx = (int)(tile % L.ntx);
    const int ty = (int)((tile / L.ntx) % L.nty);
    const int tz = (int)(tile / ((int64_t)L.ntx * L.nty));
    const int64_t x0 = (int64_t)tx * TX, y0 = (int64_t)ty * TY;
    const int yb0 = ty;
    const int64_t z_lo = (int64_t)tz * L.zc, z_hi = min(z_lo + L.zc, L.Z);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const CWeights Wt = (CWeights)(a.W + L.w_off);

    double n1 = 0.0, n2 = 0.0;   /* epilogue 2 partial norms */

    /* window plane q covers z = z_lo - hz + q; q = 0 .. (z_hi - z_lo) + 2 hz - 1 */
    const int64_t nq = (z_hi - z_lo) + 2 * hz;
    for (int64_t q = 0; q < nq; ++q) {
        const int64_t zq = z_lo - hz + q;
        const int slot = (int)q % sz;
        float* V = SEP ? s.V : s.ring + (size_t)slot * WX * WYS;
        uint8_t* ylist = s.ylist + (size_t)slot * TX * TY;
        int* cnt = s.cnt + slot * TX;
        uint64_t* base = s.base + slot * TX;
        const bool zin = zq >= 0 && zq < L.Z;
        const bool plane_zero = !a.reflect && !zin;
        const int64_t zs = a.reflect ? refl_c(zq, L.Z) : zq;
        /* 1. row lookups of the window plane, one lane per row (particle rows, then tree rows):
           the particle range to load and, for interior rows, the block's particles; and the
           coarse image (or zero) under the whole window.  Every load of this step is
           independent of the others. */
        const int nmd = 2 * WX;
        for (int t = threadIdx.x; t < nmd; t += CT) {
            const bool trow = t >= WX;
            const int xw = trow ? t - WX : t;
            const int64_t x = x0 - hx + xw;
            const bool xin = x >= 0 && x < L.X;
            const bool interior = !trow && xw >= hx && xw < hx + TX;
            uint64_t lo = 0;
            int n = 0;
            if (!(plane_zero || (!a.reflect && !xin) || (trow && L.t_row_begin < 0))) {
                const int64_t xs = a.reflect ? refl_c(x, L.X) : x;
                const int64_t row = zs * L.X + xs;
                uint64_t b, e, s0, s1;
                if (!trow) row_block(a.xz_end, a.yidx + L.yidx_off, row, L.nyb, L.row_begin + row, yb0, b, e, s0, s1);
                else row_block(a.t_xz_end, a.t_yidx + L.t_yidx_off, row, L.nyb, L.t_row_begin + row, yb0, b, e, s0, s1);
                lo = (s0 - b > (uint64_t)hy) ? s0 - hy : b;
                const uint64_t hi = (e - s1 > (uint64_t)hy) ? s1 + hy : e;
                n = (int)(hi - lo);
                if (interior) {
                    cnt[xw - hx] = (zin && xin) ? (int)min(s1 - s0, (uint64_t)TY) : 0;
                    base[xw - hx] = s0;
                }
            } else if (interior) {
                cnt[xw - hx] = 0;
                base[xw - hx] = 0;
            }
            s.mlo[t] = lo;
            s.mpre[t] = n;
        }
        /* the coarse image under the window, cells numbered over (row, y): eight loads per lane
           in flight before their stores */
        {
            const int ncell = WX * WY;
            for (int c0 = threadIdx.x; c0 < ncell; c0 += 8 * CT) {
                float cv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int c = c0 + u * CT;
                    cv[u] = 0.0f;
                    if (c < ncell && a.coarse && !plane_zero) {
                        const int xw = c / WY, yw = c - xw * WY;
                        const int64_t x = x0 - hx + xw, y = y0 - hy + yw;
                        if ((a.reflect || (x >= 0 && x < L.X)) && y >= 0 && y < L.Y) {
                            const int64_t xs = a.reflect ? refl_c(x, L.X) : x;
                            cv[u] = a.coarse[((zs >> 1) * L.cX + (xs >> 1)) * L.cY + (y >> 1)];
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int c = c0 + u * CT;
                    if (c < ncell) V[(c / WY) * WYS + (c % WY)] = cv[u];
                }
            }
        }
        __syncthreads();
        /* 2. exclusive prefix of the load counts (wave 0; nmd <= 64 here, else one lane walks) */
        if (wv == 0) {
            if (nmd <= 64) {
                const int c = lane < nmd ? s.mpre[lane] : 0;
                int inc = c;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int u = __shfl_up(inc, o, 64);
                    if (lane >= o) inc += u;
                }
                if (lane < nmd) s.mpre[lane] = inc - c;
                if (lane == nmd - 1) s.mpre[nmd] = inc;
            } else if (lane == 0) {
                int acc = 0;
                for (int t = 0; t < nmd; ++t) {
                    const int c = s.mpre[t];
                    s.mpre[t] = acc;
                    acc += c;
                }
                s.mpre[nmd] = acc;
            }
        }
        __syncthreads();
        /* 3. particles and tree nodes of the level over the coarse image (the two never share a
           cell), all rows' loads at once: four entries per lane before any is stored */
        {
            const int total = s.mpre[nmd];
            for (int k0 = threadIdx.x; k0 < total; k0 += 4 * CT) {
                int rr[4];
                uint64_t ii[4];
                int yy[4];
                float vv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + u * CT;
                    int lo = 0, hi = nmd;   /* last row r with mpre[r] <= k */
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (s.mpre[mid] <= k) lo = mid;
                        else hi = mid;
                    }
                    rr[u] = k < total ? lo : -1;
                    ii[u] = k < total ? s.mlo[lo] + (uint64_t)(k - s.mpre[lo]) : 0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (rr[u] < 0) continue;
                    if (rr[u] < WX) {
                        yy[u] = a.y_vec[ii[u]];
                        vv[u] = tof(a.in[ii[u]]);
                    } else {
                        yy[u] = a.t_y_vec[ii[u]];
                        vv[u] = a.tree[ii[u]];
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (rr[u] < 0) continue;
                    const int xw = rr[u] < WX ? rr[u] : rr[u] - WX;
                    const int yw = yy[u] - (int)(y0 - hy);
                    if (yw >= 0 && yw < WY) V[xw * WYS + yw] = vv[u];
                    if (rr[u] < WX && xw >= hx && xw < hx + TX) {
                        const uint64_t d = ii[u] - base[xw - hx];
                        if (d < (uint64_t)cnt[xw - hx]) ylist[(xw - hx) * TY + (int)d] = (uint8_t)(yy[u] - (int)y0);
                    }
                }
            }
        }
        __syncthreads();
        /* 3. y ends: mirror inside the window (every mirror source of a needed halo cell lies in it) */
        if (a.reflect && (y0 - hy < 0 || y0 + TY + hy > L.Y)) {
            for (int xw = wv; xw < WX; xw += NWAVE) {
                float* vrow = V + xw * WYS;
                for (int yw = lane; yw < WY; yw += 64) {
                    const int64_t y = y0 - hy + yw;
                    if ((y < 0 || y >= L.Y) && y < L.Y + hy) {
                        const int64_t src = refl_c(y, L.Y) - (y0 - hy);
                        if (src >= 0 && src < WY) vrow[yw] = vrow[src];
                    }
                }
            }
            __syncthreads();
        }
        /* the assembled interior is the next level's coarse image */
        if (a.vout && zq >= z_lo && zq < z_hi) {
            for (int r = wv; r < TX; r += NWAVE) {
                const int64_t x = x0 + r;
                const int64_t y = y0 + lane;
                if (x < L.X && y < L.Y) a.vout[(zq * L.X + x) * L.Y + y] = V[(r + hx) * WYS + hy + lane];
            }
        }
        if (SEP) {
            /* y pass over every window row, then x pass into the ring */
            const CWeights fy = Wt;
            const CWeights fx = Wt + L.sy;
            for (int xw = wv; xw < WX; xw += NWAVE) {
                const float* vrow = s.V + xw * WYS + lane;
                float acc = 0.0f;
                for (int k = 0; k < L.sy; ++k) acc += fy[k] * vrow[k];
                s.A[xw * TY + lane] = acc;
            }
            __syncthreads();
            float* B = s.ring + (size_t)slot * TX * TY;
            for (int r = wv; r < TX; r += NWAVE) {
                float acc = 0.0f;
                for (int k = 0; k < L.sx; ++k) acc += fx[k] * s.A[(r + k) * TY + lane];
                B[r * TY + lane] = acc;
            }
            __syncthreads();
        }
        /* 4. outputs of plane zo = zq - hz once its neighbourhood is in: a row's particles in
           the 64-cell block are at most 64, one per lane; the taps run in wave-uniform control
           flow (inactive lanes compute on cell 0 and store nothing) */
        if (q >= 2 * hz) {
            const int64_t qo = q - 2 * hz;           /* window plane of z = zo - hz */
            const int oslot = ((int)qo + hz) % sz;   /* plane zo itself (32-bit: q < Z + 2 hz) */
            /* the plane's particles numbered over the tile's rows (row r holds cnt[r]), one lane
               each; SEP keeps a wave per row */
            int ntot = 0;
            if (!SEP)
                for (int r = 0; r < TX; ++r) ntot += s.cnt[oslot * TX + r];
            for (int it = SEP ? wv : threadIdx.x; SEP ? it < TX : it - lane < ntot; it += SEP ? NWAVE : CT) {
                int r, d;
                bool act;
                if (SEP) {
                    r = it;
                    d = lane;
                    act = lane < s.cnt[oslot * TX + r];
                } else {
                    /* row of particle it: the last r with (sum of cnt below r) <= it */
                    int k = it, rr = 0;
                    act = it < ntot;
                    if (act)
                        for (; rr < TX - 1; ++rr) {
                            const int c = s.cnt[oslot * TX + rr];
                            if (k < c) break;
                            k -= c;
                        }
                    r = rr;
                    d = act ? k : 0;
                }
                if (SEP && __builtin_amdgcn_readfirstlane(s.cnt[oslot * TX + r]) == 0) continue;
                const uint64_t pb = s.base[oslot * TX + r] + (uint64_t)d;
                const int yo = act ? (int)s.ylist[((size_t)oslot * TX + r) * TY + d] : 0;
                float val = 0.0f;
                if (SEP) {
                    const CWeights fz = Wt + L.sy + L.sx;
                    for (int iz = 0; iz < sz; ++iz) {
                        const int sl = (int)((qo + iz) % sz);
                        val += fz[iz] * s.ring[((size_t)sl * TX + r) * TY + yo];
                    }
                } else {
                    CWeights w = Wt;
                    for (int iz = 0; iz < sz; ++iz) {
                        const float* pl = s.ring + (size_t)(((int)qo + iz) % sz) * WX * WYS + r * WYS + yo;
                        for (int ix = 0; ix < L.sx; ++ix) {
                            val = row_taps<SY>(pl + ix * WYS, w, L.sy, val);
                            w += L.sy;
                        }
                    }
                }
                if (act) {
                    const uint64_t i = pb;
                    if (a.epi == 0) {
                        a.out[i] = val;
                    } else if (a.epi == 1) {
                        const float num = a.epi_in_u16 ? (float)((const uint16_t*)a.epi_in)[i] : ((const float*)a.epi_in)[i];
                        a.out[i] = num / val;
                    } else {
                        const float uo = a.out[i];
                        const float nu = uo * val;
                        const double d = (double)nu - (double)uo;
                        n1 += d * d;
                        n2 += (double)uo * (double)uo;
                        a.out[i] = nu;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (a.epi == 2) {
        /* block partials of the norms, reduced in a fixed order */
        double* red = (double*)smem;
        red[threadIdx.x] = n1;
        red[CT + threadIdx.x] = n2;
        __syncthreads();
        for (int o = CT / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                red[threadIdx.x] += red[threadIdx.x + o];
                red[CT + threadIdx.x] += red[CT + threadIdx.x + o];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            a.part[2 * (int64_t)blockIdx.x] = red[0];
            a.part[2 * (int64_t)blockIdx.x + 1] = red[CT];
        }
    }
}

/* smallest power of two above n */
constexpr int pow2_above(int n) {
    int p = 1;
    while (p <= n) p <<= 1;
    return p;
}

/* Separable mode with the z pass in registers.  Each thread owns one y (its lane) in TX / 4
   rows of the tile and keeps the x-passed values of the last SZ planes of those cells in
   registers; a per-plane map from interior cells to their particle (offset in the row's block,
   0xFF for none) sits in an LDS ring of SZ + 1 planes, so the tile needs no LDS ring of planes
   and more blocks fit on a CU.  Per plane: assemble the window (as k_conv_tile), y pass into A,
   x pass into the registers, and the outputs of plane z - hz at the cells that hold particles. */
template <typename TIn, int TX, int SZ, int EPI, bool TOP>
__global__ __launch_bounds__(CT) void k_conv_sep(ConvArgs<TIn> a) {
    extern __shared__ __align__(16) uint8_t smem[];
    constexpr int SZR = SZ + 1;
    constexpr int HZ = (SZ - 1) / 2;
    constexpr int RPT = TX / NWAVE;
    const LevelDesc& L = a.L;
    constexpr int hy = HZ, hx = HZ;
    constexpr int WX = TX + 2 * hx, WY = TY + 2 * hy, WYS = WY;
    uint8_t* p = smem;
    uint64_t* sbase = (uint64_t*)p;      /* SZR x TX: first particle of each interior row's block */
    p += (size_t)SZR * TX * 8;
    uint64_t* mlo = (uint64_t*)p;        /* 2 WX */
    p += (size_t)2 * WX * 8;
    /* exclusive prefix of the window rows' load counts, then the total, then INT_MAX up to a
       power of two: the row search runs a fixed number of steps */
    constexpr int NPM = pow2_above(2 * WX);
    int* mpre = (int*)p;                 /* NPM + 2 */
    p += (size_t)(NPM + 2) * 4;
    float* V = (float*)p;                /* WX x WYS */
    p += (size_t)WX * WYS * 4;
    float* Ab = (float*)p;               /* WX x TY */
    p += (size_t)WX * TY * 4;
    int32_t* cxo = (int32_t*)p;          /* WX: coarse-plane offset of each window row, -1 none */
    p += (size_t)((WX + 3) & ~3) * 4;
    uint8_t* pidx = p;                   /* SZR x TX x TY */

    const int64_t ntiles = (int64_t)L.ntx * L.nty * L.ntz;
    const int64_t tile = xcd_block(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    /* y tiles fastest: consecutive blocks (one XCD's run) cover whole rows, so the partial
       cache lines of a row's particle chunks meet in one L2 */
    const int ty = (int)(tile % L.nty);
    const int tx = (int)((tile / L.nty) % L.ntx);
    const int tz = (int)(tile / ((int64_t)L.ntx * L.nty));
    const int64_t x0 = (int64_t)tx * TX, y0 = (int64_t)ty * TY;
    const int yb0 = ty;
    const int64_t z_lo = (int64_t)tz * L.zc, z_hi = min(z_lo + L.zc, L.Z);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float fy[SZ], fx[SZ], fz[SZ];
    {
        const CWeights w = (CWeights)(a.W + L.w_off);
#pragma unroll
        for (int k = 0; k < SZ; ++k) {
            fy[k] = w[k];
            fx[k] = w[SZ + k];
            fz[k] = w[2 * SZ + k];
        }
    }
    constexpr int nmd = TOP ? WX : 2 * WX;   /* window rows looked up: particle rows, tree rows */
    constexpr int NP = pow2_above(nmd);
    for (int t = nmd + 1 + threadIdx.x; t < NP; t += CT) mpre[t] = 0x7FFFFFFF;

    /* plane-invariant coarse geometry: each window row's offset in a coarse plane (LDS), and
       this lane's two window columns' coarse y (-1 outside the level) */
    for (int t = threadIdx.x; t < WX; t += CT) {
        const int64_t x = x0 - hx + t;
        const bool ok = a.coarse && (a.reflect || (x >= 0 && x < L.X));
        cxo[t] = ok ? (int32_t)(((a.reflect ? refl_c(x, L.X) : x) >> 1) * L.cY) : -1;
    }
    int32_t cyo[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int yw = lane + 64 * h;
        const int64_t y = y0 - hy + yw;
        cyo[h] = (yw < WY && y >= 0 && y < L.Y) ? (int32_t)(y >> 1) : -1;
    }
    const int64_t cplane = L.cX * L.cY;
    __syncthreads();

    double n1 = 0.0, n2 = 0.0;
    float breg[RPT][SZ];
#pragma unroll
    for (int j = 0; j < RPT; ++j)
#pragma unroll
        for (int k = 0; k < SZ; ++k) breg[j][k] = 0.0f;

    const int64_t nq = (z_hi - z_lo) + 2 * HZ;
    for (int64_t q = 0; q < nq; ++q) {
        const int64_t zq = z_lo - HZ + q;
        const int slot = (int)(q % SZR);
        uint8_t* pid = pidx + (size_t)slot * TX * TY;
        uint64_t* bs = sbase + (size_t)slot * TX;
        const bool zin = zq >= 0 && zq < L.Z;
        const bool plane_zero = !a.reflect && !zin;
        const int64_t zs = a.reflect ? refl_c(zq, L.Z) : zq;
        /* 1. wave 0: row lookups and the exclusive prefix of the load counts; every thread:
           clear this plane's particle map; all waves: the coarse image under the window */
        for (int c = threadIdx.x * 8; c < TX * TY; c += CT * 8) *(uint64_t*)(pid + c) = ~(uint64_t)0;
        if (wv == 0) {
            int carry = 0;
            for (int t0 = 0; t0 < nmd; t0 += 64) {
                const int t = t0 + lane;
                uint64_t lo = 0;
                int n = 0;
                if (t < nmd) {
                    const bool trow = t >= WX;
                    const int xw = trow ? t - WX : t;
                    const int64_t x = x0 - hx + xw;
                    const bool xin = x >= 0 && x < L.X;
                    const bool interior = !trow && xw >= hx && xw < hx + TX;
                    if (!(plane_zero || (!a.reflect && !xin) || (trow && (TOP || L.t_row_begin < 0)))) {
                        const int64_t xs = a.reflect ? refl_c(x, L.X) : x;
                        const int64_t row = zs * L.X + xs;
                        uint64_t b, e, s0, s1;
                        if (TOP || !trow) row_block(a.xz_end, a.yidx + L.yidx_off, row, L.nyb, L.row_begin + row, yb0, b, e, s0, s1);
                        else row_block(a.t_xz_end, a.t_yidx + L.t_yidx_off, row, L.nyb, L.t_row_begin + row, yb0, b, e, s0, s1);
                        lo = (s0 - b > (uint64_t)hy) ? s0 - hy : b;
                        const uint64_t hi = (e - s1 > (uint64_t)hy) ? s1 + hy : e;
                        n = (int)(hi - lo);
                        /* interior rows of planes inside the level list their block's particles */
                        if (interior) bs[xw - hx] = (zin && xin) ? s0 : ~(uint64_t)0;
                    } else if (interior) {
                        bs[xw - hx] = ~(uint64_t)0;
                    }
                    mlo[t] = lo;
                }
                int inc = n;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int u = __shfl_up(inc, o, 64);
                    if (lane >= o) inc += u;
                }
                if (t < nmd) mpre[t] = carry + inc - n;
                carry += __shfl(inc, 63, 64);
            }
            if (lane == 0) mpre[nmd] = carry;
        }
        /* the coarse image under the window: a wave per window row, lanes along y (a second
           pass for the y halo beyond 64 cells), four rows' loads in flight before their stores;
           offsets within the coarse plane come from the plane-invariant tables */
        const float* cpl = (plane_zero || !a.coarse) ? nullptr : a.coarse + (zs >> 1) * cplane;
        for (int xb = wv; xb < WX; xb += 4 * NWAVE) {
            float cv[4][2];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int xw = xb + u * NWAVE;
                const int32_t xo = (cpl && xw < WX) ? cxo[xw] : -1;
#pragma unroll
                for (int h = 0; h < 2; ++h) cv[u][h] = (xo >= 0 && cyo[h] >= 0) ? cpl[xo + cyo[h]] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int xw = xb + u * NWAVE;
                if (xw >= WX) break;
                V[xw * WYS + lane] = cv[u][0];
                if (lane + 64 < WY) V[xw * WYS + lane + 64] = cv[u][1];
            }
        }
        __syncthreads();
        /* 2. particles and tree nodes of the level over the coarse image; interior particles of
           planes inside the level enter the plane's particle map */
        {
            const int total = mpre[nmd];
            for (int k0 = threadIdx.x; k0 < total; k0 += 4 * CT) {
                int rr[4];
                uint64_t ii[4];
                int yy[4];
                float vv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + u * CT;
                    /* the last row r with mpre[r] <= k, in log2(NP) fixed steps */
                    int lo = 0;
#pragma unroll
                    for (int st = NP / 2; st >= 1; st >>= 1)
                        if (mpre[lo + st] <= k) lo += st;
                    rr[u] = k < total ? lo : -1;
                    ii[u] = k < total ? mlo[lo] + (uint64_t)(k - mpre[lo]) : 0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (rr[u] < 0) continue;
                    if (TOP || rr[u] < WX) {
                        yy[u] = a.y_vec[ii[u]];
                        vv[u] = tof(a.in[ii[u]]);
                    } else {
                        yy[u] = a.t_y_vec[ii[u]];
                        vv[u] = a.tree[ii[u]];
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (rr[u] < 0) continue;
                    const int xw = rr[u] < WX ? rr[u] : rr[u] - WX;
                    const int yw = yy[u] - (int)(y0 - hy);
                    if (yw >= 0 && yw < WY) V[xw * WYS + yw] = vv[u];
                    if (rr[u] < WX && xw >= hx && xw < hx + TX) {
                        const uint64_t b0 = bs[xw - hx];
                        const uint64_t d = ii[u] - b0;
                        const int yo = yy[u] - (int)y0;
                        if (b0 != ~(uint64_t)0 && d < (uint64_t)TY && yo >= 0 && yo < TY) pid[(xw - hx) * TY + yo] = (uint8_t)d;
                    }
                }
            }
        }
        __syncthreads();
        /* 3. y ends: mirror inside the window */
        if (a.reflect && (y0 - hy < 0 || y0 + TY + hy > L.Y)) {
            for (int xw = wv; xw < WX; xw += NWAVE) {
                float* vrow = V + xw * WYS;
                for (int yw = lane; yw < WY; yw += 64) {
                    const int64_t y = y0 - hy + yw;
                    if ((y < 0 || y >= L.Y) && y < L.Y + hy) {
                        const int64_t src = refl_c(y, L.Y) - (y0 - hy);
                        if (src >= 0 && src < WY) vrow[yw] = vrow[src];
                    }
                }
            }
            __syncthreads();
        }
        /* the assembled interior is the next level's coarse image */
        if (!TOP && a.vout && zq >= z_lo && zq < z_hi) {
            for (int r = wv; r < TX; r += NWAVE) {
                const int64_t x = x0 + r;
                const int64_t y = y0 + lane;
                if (x < L.X && y < L.Y) a.vout[(zq * L.X + x) * L.Y + y] = V[(r + hx) * WYS + hy + lane];
            }
        }
        /* 4. y pass of every window row */
        for (int xw = wv; xw < WX; xw += NWAVE) {
            const float* vrow = V + xw * WYS + lane;
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < SZ; ++k) acc = fmaf(fy[k], vrow[k], acc);
            Ab[xw * TY + lane] = acc;
        }
        __syncthreads();
        /* 5. x pass of the thread's rows into the register ring; then the outputs of plane
           zq - HZ, whose z neighbourhood is now in the ring */
        const bool emit = q >= 2 * HZ;
        const int oslot = (int)((q - HZ + SZR) % SZR);
        const uint8_t* opid = pidx + (size_t)oslot * TX * TY;
        const uint64_t* obs = sbase + (size_t)oslot * TX;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int r = wv + j * NWAVE;
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < SZ; ++k) acc = fmaf(fx[k], Ab[(r + k) * TY + lane], acc);
#pragma unroll
            for (int k = 0; k + 1 < SZ; ++k) breg[j][k] = breg[j][k + 1];
            breg[j][SZ - 1] = acc;
        }
        if (emit) {
            uint64_t oi[RPT];
            float ov[RPT];
            bool on[RPT];
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                const int r = wv + j * NWAVE;
                const int d = opid[r * TY + lane];
                on[j] = d != 0xFF;
                oi[j] = on[j] ? obs[r] + (uint64_t)d : 0;
                float val = 0.0f;
                if (on[j]) {
#pragma unroll
                    for (int k = 0; k < SZ; ++k) val = fmaf(fz[k], breg[j][k], val);
                }
                ov[j] = val;
            }
            if constexpr (EPI == 0) {
#pragma unroll
                for (int j = 0; j < RPT; ++j)
                    if (on[j]) a.out[oi[j]] = ov[j];
            } else if constexpr (EPI == 1) {
                float num[RPT];
#pragma unroll
                for (int j = 0; j < RPT; ++j)
                    num[j] = on[j] ? (a.epi_in_u16 ? (float)((const uint16_t*)a.epi_in)[oi[j]] : ((const float*)a.epi_in)[oi[j]]) : 1.0f;
#pragma unroll
                for (int j = 0; j < RPT; ++j)
                    if (on[j]) a.out[oi[j]] = num[j] / ov[j];
            } else {
                float uo[RPT];
#pragma unroll
                for (int j = 0; j < RPT; ++j) uo[j] = on[j] ? a.out[oi[j]] : 0.0f;
#pragma unroll
                for (int j = 0; j < RPT; ++j)
                    if (on[j]) {
                        const float nu = uo[j] * ov[j];
                        const double d = (double)nu - (double)uo[j];
                        n1 += d * d;
                        n2 += (double)uo[j] * (double)uo[j];
                        a.out[oi[j]] = nu;
                    }
            }
        }
    }
    if constexpr (EPI == 2) {
        __syncthreads();
        double* red = (double*)smem;
        red[threadIdx.x] = n1;
        red[CT + threadIdx.x] = n2;
        __syncthreads();
        for (int o = CT / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                red[threadIdx.x] += red[threadIdx.x + o];
                red[CT + threadIdx.x] += red[CT + threadIdx.x + o];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            a.part[2 * (int64_t)blockIdx.x] = red[0];
            a.part[2 * (int64_t)blockIdx.x + 1] = red[CT];
        }
    }
}

size_t conv_sep_lds_bytes(int tx, int sy, int sx, int sz) {
    const int WX = tx + (sx - 1), WYS = TY + (sy - 1), SZR = sz + 1;
    size_t b = (size_t)SZR * tx * 8 + (size_t)2 * WX * 8 + (size_t)(pow2_above(2 * WX) + 2) * 4 + (size_t)WX * WYS * 4 + (size_t)WX * TY * 4 +
               (size_t)((WX + 3) & ~3) * 4 + (size_t)SZR * tx * TY;
    return std::max<size_t>(b, 2 * CT * sizeof(double));
}

size_t conv_lds_bytes(int tx, bool sep, int sy, int sx, int sz) {
    const int WX = tx + (sx - 1), WYS = TY + (sy - 1);
    size_t b = sep ? ((size_t)sz * tx * TY + (size_t)WX * WYS + (size_t)WX * TY) * 4 : (size_t)sz * WX * WYS * 4;
    b += (size_t)sz * tx * (8 + 4 + TY);
    b += (size_t)2 * WX * 8 + (size_t)(2 * WX + 2) * 4;
    return std::max<size_t>(b, 2 * CT * sizeof(double));
}

/* y-block index of one level (or tree level): for rows with nyb >= 2, entry (row, yb - 1),
   yb = 1 .. nyb-1, is the number of the row's particles with y < 64 yb.  One wave per row. */
__global__ __launch_bounds__(256) void k_yidx(const uint64_t* __restrict__ xz, const uint16_t* __restrict__ y_vec, int64_t row_begin,
                                              int64_t nrows, int nyb, uint16_t* idx) {
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / 64;
    for (int64_t r = wid; r < nrows; r += nw) {
        const uint64_t b = xz[row_begin + r - 1], e = xz[row_begin + r];
        uint16_t* ix = idx + r * (nyb - 1);
        if (e == b) {
            for (int yb = 1 + lane; yb < nyb; yb += 64) ix[yb - 1] = 0;
            continue;
        }
        for (uint64_t i = b + lane; i < e; i += 64) {
            const int cb = (int)y_vec[i] >> 6;
            const int from = (i == b) ? 1 : ((int)y_vec[i - 1] >> 6) + 1;
            const int to = min(cb, nyb - 1);
            for (int yb = max(from, 1); yb <= to; ++yb) ix[yb - 1] = (uint16_t)(i - b);
            if (i + 1 == e)
                for (int yb = max(cb + 1, 1); yb < nyb; ++yb) ix[yb - 1] = (uint16_t)(e - b);
        }
    }
}

/* the top tree level (children: the particles of l_max; see the file comment).  check_only:
   set *bad when a row does not have the shape the direct reading assumes. */
template <typename T>
__global__ __launch_bounds__(256) void k_tree_top(const uint64_t* __restrict__ xz, const T* __restrict__ in, const uint64_t* __restrict__ txz,
                                                  const uint16_t* __restrict__ ty, int64_t t_row_begin, int64_t c_row_begin, int64_t pyn,
                                                  int64_t pxn, int64_t pzn, int64_t cyn, int64_t cxn, int64_t czn, float* tree,
                                                  int check_only, int* bad) {
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / 64;
    const int64_t nrows = pxn * pzn;
    for (int64_t row = wid; row < nrows; row += nw) {
        const int64_t zp = row / pxn, xp = row % pxn;
        const uint64_t tb = txz[t_row_begin + row - 1], te = txz[t_row_begin + row];
        const int64_t n = (int64_t)(te - tb);
        const int nz = (int)min((int64_t)2, czn - 2 * zp), nx = (int)min((int64_t)2, cxn - 2 * xp);
        const int64_t expect = 2 * n - ((n > 0 && (int64_t)ty[te - 1] == pyn - 1 && 2 * pyn != cyn) ? 1 : 0);
        uint64_t cb[2][2] = {{0, 0}, {0, 0}};
        bool ok = true;
        for (int za = 0; za < nz; ++za)
            for (int xb = 0; xb < nx; ++xb) {
                const int64_t crow = c_row_begin + (2 * zp + za) * cxn + (2 * xp + xb);
                cb[za][xb] = xz[crow - 1];
                if ((int64_t)(xz[crow] - cb[za][xb]) != expect) ok = false;
            }
        if (check_only) {
            if (!ok && lane == 0) atomicOr(bad, 1);
            continue;
        }
        /* boundary scale factors (APRTreeNumerics.hpp:99-113) */
        float sxz = (float)((((2 * pxn != cxn) && (xp == pxn - 1)) + ((2 * pzn != czn) && (zp == pzn - 1))) * 2);
        if (sxz == 0) sxz = 1;
        const float syxz = (2 * pyn != cyn) ? sxz * 2 : sxz;
        for (int64_t k = lane; k < n; k += 64) {
            const int64_t yp = ty[tb + k];
            const float sc = (yp == pyn - 1) ? syxz : sxz;
            const int ny = (int)min((int64_t)2, cyn - 2 * yp);
            float acc = 0.0f;
            for (int za = 0; za < nz; ++za)
                for (int xb = 0; xb < nx; ++xb)
                    for (int c = 0; c < ny; ++c) acc = sc * (float)in[cb[za][xb] + 2 * k + c] / 8.0f + acc;
            tree[tb + k] = acc;
        }
    }
}

/* sums of the per-block norm partials in a fixed order: out[0], out[1] */
__global__ __launch_bounds__(256) void k_reduce_parts(const double* __restrict__ part, int64_t nb, double* out) {
    __shared__ double r1[256], r2[256];
    double a = 0, b = 0;
    for (int64_t i = threadIdx.x; i < nb; i += 256) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    r1[threadIdx.x] = a;
    r2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            r1[threadIdx.x] += r1[threadIdx.x + o];
            r2[threadIdx.x] += r2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = r1[0];
        out[1] = r2[0];
    }
}

int nyb_of(int64_t y) { return (int)cdiv(y, TY); }

}  // namespace

void ensure_yidx(aprh_context* ctx, aprh_apr& a) {
    if (a.yidx_built) return;
    ensure_tree(ctx, a);
    hipStream_t st = ctx->stream;
    a.yidx_off.assign(a.l_max + 1, 0);
    a.t_yidx_off.assign(a.t_l_max + 1, 0);
    int64_t na = 0, nt = 0;
    for (int l = a.l_min; l <= a.l_max; ++l) {
        a.yidx_off[l] = na;
        const int nyb = nyb_of(a.y_num[l]);
        if (nyb >= 2) na += a.x_num[l] * a.z_num[l] * (nyb - 1);
    }
    for (int t = 0; t <= a.t_l_max; ++t) {
        a.t_yidx_off[t] = nt;
        const int nyb = nyb_of(a.t_y[t]);
        if (nyb >= 2) nt += a.t_x[t] * a.t_z[t] * (nyb - 1);
    }
    a.yidx.reset(ctx->pool, (size_t)std::max<int64_t>(na, 1) * 2);
    a.t_yidx.reset(ctx->pool, (size_t)std::max<int64_t>(nt, 1) * 2);
    for (int l = a.l_min; l <= a.l_max; ++l) {
        const int nyb = nyb_of(a.y_num[l]);
        const int64_t nrows = a.x_num[l] * a.z_num[l];
        if (nyb < 2 || nrows == 0) continue;
        hipLaunchKernelGGL(k_yidx, dim3((unsigned)std::min<int64_t>(cdiv(nrows, 4), 16384)), dim3(256), 0, st, a.xz_end.as<uint64_t>(),
                           a.y_vec.as<uint16_t>(), (int64_t)a.level_xz_vec[l], nrows, nyb, a.yidx.as<uint16_t>() + a.yidx_off[l]);
        HIP_LAUNCH_CHECK();
    }
    for (int t = 0; t <= a.t_l_max; ++t) {
        const int nyb = nyb_of(a.t_y[t]);
        const int64_t nrows = a.t_x[t] * a.t_z[t];
        if (nyb < 2 || nrows == 0 || a.t_total == 0) continue;
        hipLaunchKernelGGL(k_yidx, dim3((unsigned)std::min<int64_t>(cdiv(nrows, 4), 16384)), dim3(256), 0, st, a.t_xz_end.as<uint64_t>(),
                           a.t_y_vec.as<uint16_t>(), (int64_t)a.t_level_xz[t], nrows, nyb, a.t_yidx.as<uint16_t>() + a.t_yidx_off[t]);
        HIP_LAUNCH_CHECK();
    }
    HIP_CHECK(hipStreamSynchronize(st));
    a.yidx_built = true;
}

/* the top tree level's means straight from the particle rows; false when this APR's rows do
   not have that shape (the caller then takes the general path) */
template <typename T>
bool tree_mean_top(aprh_context* ctx, aprh_apr& a, const T* parts, float* tree) {
    hipStream_t st = ctx->stream;
    const int t = a.t_l_max, c = t + 1;
    if (c != a.l_max || c < a.l_min || a.t_total == 0) return false;
    const int64_t pyn = a.t_y[t], pxn = a.t_x[t], pzn = a.t_z[t];
    const int64_t nrows = pxn * pzn;
    const unsigned grid = (unsigned)std::min<int64_t>(std::max<int64_t>(cdiv(nrows, 4), 1), 16384);
    if (a.top_tree_regular < 0) {
        int* bad = ctx->d_flags + 32;
        HIP_CHECK(hipMemsetAsync(bad, 0, 4, st));
        hipLaunchKernelGGL((k_tree_top<T>), dim3(grid), dim3(256), 0, st, a.xz_end.as<uint64_t>(), parts, a.t_xz_end.as<uint64_t>(),
                           a.t_y_vec.as<uint16_t>(), (int64_t)a.t_level_xz[t], (int64_t)a.level_xz_vec[c], pyn, pxn, pzn, a.y_num[c],
                           a.x_num[c], a.z_num[c], tree, 1, bad);
        HIP_LAUNCH_CHECK();
        int h = 0;
        HIP_CHECK(hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        a.top_tree_regular = h ? 0 : 1;
    }
    if (!a.top_tree_regular) return false;
    hipLaunchKernelGGL((k_tree_top<T>), dim3(grid), dim3(256), 0, st, a.xz_end.as<uint64_t>(), parts, a.t_xz_end.as<uint64_t>(),
                       a.t_y_vec.as<uint16_t>(), (int64_t)a.t_level_xz[t], (int64_t)a.level_xz_vec[c], pyn, pxn, pzn, a.y_num[c],
                       a.x_num[c], a.z_num[c], tree, 0, (int*)nullptr);
    HIP_LAUNCH_CHECK();
    return true;
}

namespace {
template <typename TIn, int TX, bool SEP, int SY>
void launch_conv_tile(const ConvArgs<TIn>& args, unsigned grid, size_t lds, hipStream_t st) {
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)k_conv_tile<TIn, TX, SEP, SY>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_conv_tile<TIn, TX, SEP, SY>), dim3(grid), dim3(CT), lds, st, args);
    HIP_LAUNCH_CHECK();
}

template <typename TIn, int TX, int SZ, int EPI, bool TOP>
void launch_conv_sep_k(const ConvArgs<TIn>& args, unsigned grid, size_t lds, hipStream_t st) {
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)k_conv_sep<TIn, TX, SZ, EPI, TOP>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_conv_sep<TIn, TX, SZ, EPI, TOP>), dim3(grid), dim3(CT), lds, st, args);
    HIP_LAUNCH_CHECK();
}

/* the variant for the launch's epilogue and level: the finest level has no tree nodes and no
   next-level image; u16 particles only ever feed the plain store */
template <typename TIn, int TX, int SZ>
void launch_conv_sep_tile(const ConvArgs<TIn>& a, unsigned grid, size_t lds, hipStream_t st) {
    const bool top = a.L.t_row_begin < 0 && a.vout == nullptr;
    if constexpr (std::is_same<TIn, uint16_t>::value) {
        if (a.epi != 0) throw Error(APRH_ERR_UNSUPPORTED, "convolution epilogue needs float particles");
        if (top) launch_conv_sep_k<TIn, TX, SZ, 0, true>(a, grid, lds, st);
        else launch_conv_sep_k<TIn, TX, SZ, 0, false>(a, grid, lds, st);
    } else {
        switch (a.epi * 2 + (top ? 1 : 0)) {
            case 0: launch_conv_sep_k<TIn, TX, SZ, 0, false>(a, grid, lds, st); break;
            case 1: launch_conv_sep_k<TIn, TX, SZ, 0, true>(a, grid, lds, st); break;
            case 2: launch_conv_sep_k<TIn, TX, SZ, 1, false>(a, grid, lds, st); break;
            case 3: launch_conv_sep_k<TIn, TX, SZ, 1, true>(a, grid, lds, st); break;
            case 4: launch_conv_sep_k<TIn, TX, SZ, 2, false>(a, grid, lds, st); break;
            default: launch_conv_sep_k<TIn, TX, SZ, 2, true>(a, grid, lds, st); break;
        }
    }
}

template <typename TIn, int TX>
void launch_conv_sep_sz(const ConvArgs<TIn>& ca, unsigned grid, size_t lds, hipStream_t st) {
    switch (ca.L.sz) {
        case 1: launch_conv_sep_tile<TIn, TX, 1>(ca, grid, lds, st); return;
        case 3: launch_conv_sep_tile<TIn, TX, 3>(ca, grid, lds, st); return;
        case 5: launch_conv_sep_tile<TIn, TX, 5>(ca, grid, lds, st); return;
        case 7: launch_conv_sep_tile<TIn, TX, 7>(ca, grid, lds, st); return;
        case 9: launch_conv_sep_tile<TIn, TX, 9>(ca, grid, lds, st); return;
        case 11: launch_conv_sep_tile<TIn, TX, 11>(ca, grid, lds, st); return;
        case 13: launch_conv_sep_tile<TIn, TX, 13>(ca, grid, lds, st); return;
        default: throw Error(APRH_ERR_UNSUPPORTED, "register z ring: stencil depth not instantiated");
    }
}

/* z sizes with a register-ring separable kernel */
bool sep_regz_size(int sz) { return sz >= 1 && sz <= 13 && (sz & 1); }

/* the kernel for one level: separable or reference order, tile width, unrolled y taps */
template <typename TIn>
void launch_conv_level(const ConvArgs<TIn>& ca, int tx, bool sep, unsigned grid, size_t lds, hipStream_t st) {
    if (sep) {
        if (tx == 16) launch_conv_tile<TIn, 16, true, 0>(ca, grid, lds, st);
        else launch_conv_tile<TIn, 8, true, 0>(ca, grid, lds, st);
        return;
    }
    if (tx == 16) {
        switch (ca.L.sy) {
            case 1: launch_conv_tile<TIn, 16, false, 1>(ca, grid, lds, st); return;
            case 3: launch_conv_tile<TIn, 16, false, 3>(ca, grid, lds, st); return;
            case 5: launch_conv_tile<TIn, 16, false, 5>(ca, grid, lds, st); return;
            case 7: launch_conv_tile<TIn, 16, false, 7>(ca, grid, lds, st); return;
            case 9: launch_conv_tile<TIn, 16, false, 9>(ca, grid, lds, st); return;
            default: launch_conv_tile<TIn, 16, false, 0>(ca, grid, lds, st); return;
        }
    }
    launch_conv_tile<TIn, 8, false, 0>(ca, grid, lds, st);
}
}  // namespace

/* APR convolution of `in` (tree values `tree` at the tree nodes) over every level; see the
   file comment.  stencils[min(l_max - l, n - 1)] at level l.  sep: evaluate rank-1 stencils as
   three 1-D passes.  epi: 0 out = result; 1 out = epi_in / result; 2 out *= result, with the
   block partials of the relative-update norms appended to *part (grown as needed) */
template <typename T>
void conv_levels(aprh_context* ctx, aprh_apr& a, const std::vector<HostStencil>& stencils, bool sep_mode, const T* in, const float* tree,
                 float* out, bool reflect, int epi, const void* epi_in, bool epi_u16, Buf* part, int64_t* nparts) {
    hipStream_t st = ctx->stream;
    ensure_yidx(ctx, a);
    /* weights of every stencil: reference order, and 1-D factors where separable */
    std::vector<float> flat;
    std::vector<int64_t> off_exact, off_sep;
    for (const auto& s : stencils) {
        if (s.sy % 2 == 0 || s.sx % 2 == 0 || s.sz % 2 == 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "stencil sizes must be odd");
        if (s.sy > 2 * TY + 1) throw Error(APRH_ERR_UNSUPPORTED, "stencil wider than 129 along y");
        off_exact.push_back((int64_t)flat.size());
        flat.insert(flat.end(), s.w.begin(), s.w.end());
        std::vector<float> fy, fx, fz;
        if (sep_mode && rank1_factors(s, fy, fx, fz)) {
            off_sep.push_back((int64_t)flat.size());
            flat.insert(flat.end(), fy.begin(), fy.end());
            flat.insert(flat.end(), fx.begin(), fx.end());
            flat.insert(flat.end(), fz.begin(), fz.end());
        } else {
            off_sep.push_back(-1);
        }
    }
    Buf W(ctx->pool, flat.size() * 4);
    HIP_CHECK(hipMemcpyAsync(W.ptr(), flat.data(), flat.size() * 4, hipMemcpyHostToDevice, st));
    /* images of consecutive levels (ping-pong): level l's launch reads V(l-1), writes V(l) */
    const size_t vmax = a.l_max > a.l_min ? (size_t)(a.y_num[a.l_max - 1] * a.x_num[a.l_max - 1] * a.z_num[a.l_max - 1]) : 1;
    Buf V0(ctx->pool, vmax * 4), V1(ctx->pool, a.l_max > a.l_min + 1 ? vmax * 4 : 4);
    float* vb[2] = {V0.as<float>(), V1.as<float>()};
    /* launch shapes, to size the partial-norm buffer first */
    struct Plan {
        LevelDesc L;
        int tx;
        bool sep, regz;
        size_t lds;
        unsigned grid;
    };
    std::vector<Plan> plans;
    int64_t total_blocks = 0;
    for (int l = a.l_min; l <= a.l_max; ++l) {
        const int sn = std::min((int)stencils.size() - 1, a.l_max - l);
        const HostStencil& s = stencils[(size_t)sn];
        Plan pl;
        LevelDesc& L = pl.L;
        std::memset(&L, 0, sizeof(L));
        L.Y = a.y_num[l];
        L.X = a.x_num[l];
        L.Z = a.z_num[l];
        L.row_begin = (int64_t)a.level_xz_vec[l];
        L.t_row_begin = (l <= a.t_l_max && a.t_total > 0) ? (int64_t)a.t_level_xz[l] : -1;
        L.cY = l > a.l_min ? a.y_num[l - 1] : 0;
        L.cX = l > a.l_min ? a.x_num[l - 1] : 0;
        L.yidx_off = a.yidx_off[l];
        L.t_yidx_off = l <= a.t_l_max ? a.t_yidx_off[l] : 0;
        L.nyb = nyb_of(L.Y);
        L.sy = s.sy;
        L.sx = s.sx;
        L.sz = s.sz;
        pl.sep = off_sep[(size_t)sn] >= 0;
        L.w_off = pl.sep ? off_sep[(size_t)sn] : off_exact[(size_t)sn];
        /* the register-ring kernel addresses a coarse plane with 32-bit offsets */
        pl.regz = pl.sep && sep_regz_size(s.sz) && s.sy == s.sz && s.sx == s.sz && L.cX * L.cY <= 0x7FFFFFFF;
        if (pl.regz) {
            /* 16-row tiles: more blocks per CU hide the plane's global reads better than the
               smaller x halo of 32-row tiles saves (measured on the 9x9x9 RL workload) */
            pl.tx = 16;
            pl.lds = conv_sep_lds_bytes(pl.tx, s.sy, s.sx, s.sz);
        } else {
            pl.tx = 16;
            pl.lds = conv_lds_bytes(16, pl.sep, s.sy, s.sx, s.sz);
            if (pl.lds > 96 * 1024) {
                pl.tx = 8;
                pl.lds = conv_lds_bytes(8, pl.sep, s.sy, s.sx, s.sz);
            }
        }
        if (pl.lds > 160 * 1024) throw Error(APRH_ERR_UNSUPPORTED, "stencil too large for the tile convolution");
        L.ntx = (int)cdiv(L.X, pl.tx);
        L.nty = L.nyb;
        const int64_t nxy = (int64_t)L.ntx * L.nty;
        const int hz = (s.sz - 1) / 2;
        int64_t zc;
        if (pl.regz) {
            /* about 4096 blocks (8192 for 3-plane stencils, whose short halo makes shorter runs
               cheap: 3x3x3 passes of the 1024x1024x256 RL-TV workload 2 % faster); runs of at
               least 4 hz planes unless that leaves fewer than ~1024 */
            zc = cdiv(L.Z * nxy, hz <= 1 ? 8192 : 4096);
            const int64_t zmin = std::min<int64_t>(L.Z, std::max(8, 4 * hz));
            if (nxy * cdiv(L.Z, zmin) >= 1024) zc = std::max(zc, zmin);
            zc = std::max<int64_t>(1, std::min<int64_t>(zc, L.Z));
        } else {
            /* at most about 8192 blocks, and at least about 1024 unless runs would exceed 4 hz
               planes: a block marches its planes one after another, so a small level wants short
               runs, a large one runs long enough to amortise the 2 hz halo planes */
            const int64_t W = L.Z * nxy;
            const int64_t zmin = std::min<int64_t>(L.Z, std::max(8, 4 * hz));
            zc = std::max(cdiv(W, 8192), std::min(zmin, cdiv(W, 1024)));
            zc = std::max<int64_t>(1, std::min<int64_t>(zc, L.Z));
        }
        L.zc = (int)zc;
        L.ntz = (int)cdiv(L.Z, zc);
        const int64_t g = nxy * L.ntz;
        if (g > 0x7FFFFFFF) throw Error(APRH_ERR_UNSUPPORTED, "too many convolution tiles");
        pl.grid = (unsigned)g;
        total_blocks += g;
        plans.push_back(pl);
    }
    double* pbase = nullptr;
    if (epi == 2) {
        const int64_t need = (*nparts) + total_blocks;
        if ((int64_t)(part->bytes() / 16) < need) {
            Buf nb(ctx->pool, (size_t)need * 16);
            if (*nparts > 0) HIP_CHECK(hipMemcpyAsync(nb.ptr(), part->ptr(), (size_t)(*nparts) * 16, hipMemcpyDeviceToDevice, st));
            *part = std::move(nb);
        }
        pbase = part->as<double>() + 2 * (*nparts);
        *nparts = need;
    }
    int64_t boff = 0;
    for (int l = a.l_min; l <= a.l_max; ++l) {
        const Plan& pl = plans[(size_t)(l - a.l_min)];
        ConvArgs<T> ca;
        std::memset(&ca, 0, sizeof(ca));
        ca.L = pl.L;
        ca.xz_end = a.xz_end.as<uint64_t>();
        ca.y_vec = a.y_vec.as<uint16_t>();
        ca.in = in;
        ca.t_xz_end = a.t_xz_end.as<uint64_t>();
        ca.t_y_vec = a.t_y_vec.as<uint16_t>();
        ca.tree = tree;
        ca.yidx = a.yidx.as<uint16_t>();
        ca.t_yidx = a.t_yidx.as<uint16_t>();
        ca.coarse = l > a.l_min ? vb[(l - 1) & 1] : nullptr;
        ca.vout = l < a.l_max ? vb[l & 1] : nullptr;
        ca.W = W.as<float>();
        ca.out = out;
        ca.epi_in = epi_in;
        ca.part = pbase ? pbase + 2 * boff : nullptr;
        ca.reflect = reflect ? 1 : 0;
        ca.epi = epi;
        ca.epi_in_u16 = epi_u16 ? 1 : 0;
        boff += pl.grid;
        if (pl.grid == 0) continue;
        if (pl.regz) {
            launch_conv_sep_sz<T, 16>(ca, pl.grid, pl.lds, st);
            continue;
        }
        launch_conv_level<T>(ca, pl.tx, pl.sep, pl.grid, pl.lds, st);
    }
}

void reduce_norm_parts(aprh_context* ctx, const double* part, int64_t nparts, double* d_out) {
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, ctx->stream, part, nparts, d_out);
    HIP_LAUNCH_CHECK();
}

template bool tree_mean_top<uint16_t>(aprh_context*, aprh_apr&, const uint16_t*, float*);
template bool tree_mean_top<float>(aprh_context*, aprh_apr&, const float*, float*);
template void conv_levels<uint16_t>(aprh_context*, aprh_apr&, const std::vector<HostStencil>&, bool, const uint16_t*, const float*, float*,
                                    bool, int, const void*, bool, Buf*, int64_t*);
template void conv_levels<float>(aprh_context*, aprh_apr&, const std::vector<HostStencil>&, bool, const float*, const float*, float*,
                                 bool, int, const void*, bool, Buf*, int64_t*);

}  // namespace aprh
