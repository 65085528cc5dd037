/*
 * k_structure.hip -- pulling scheme, linear particle structure and particle
 * sampling on gfx950.
 *
 *  - level pyramid + particle-cell seeds: 2x2x2 max downsample of the level
 *    map fused with PullingScheme::fill (LocalParticleCellSet.hpp:58-91,
 *    PullingScheme.hpp:140-168).
 *  - pulling scheme: one in-place pass per level, top to bottom, fusing
 *    set_ascendant_neighbours, set_filler and fill_neighbours/fill_parent
 *    (PullingScheme.hpp:111-393).  The per-cell predicates only test states
 *    that the pass cannot create or destroy (ASCENDANT; SEED-or-PROPOGATE),
 *    so concurrent in-place updates give the sequential result.
 *  - linear structure (LinearAccess.hpp:210-524): equivalence optimisation
 *    level by level, then per-row particle counts (a wave per row, 16 cells
 *    per lane), one exclusive scan over all rows in particle order (which is
 *    xz_end_vec), and a wave-per-row fill of y_vec with ballot prefix counts.
 *  - sampling (ParticleData.hpp:221-261): mean pyramid of the image, then a
 *    row-parallel gather.
 */
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "aprh_internal.hpp"
#include "tiles.hpp"

namespace aprh {

namespace {

enum : uint8_t { EMPTY = 0, SEED = 1, BOUNDARY = 2, FILLER = 3, UPSEED = 4, ASC = 8, PROP = 15, ASCN = 16 };

struct Dims {
    int64_t y, x, z;
};

/* level map of level l from level l+1 (max of the clamped 2x2x2 block) + seeds of level l */
__device__ __forceinline__ void level_down_body(const uint8_t* __restrict__ fine, Dims fd, uint8_t* coarse, Dims cd, uint8_t* pct,
                                                int level, int mode, int64_t G, int64_t t) {
    const size_t n = (size_t)(cd.y * cd.x * cd.z);
    for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < n; i += (size_t)G * blockDim.x) {
        const int64_t y = (int64_t)(i % (size_t)cd.y);
        const int64_t r = (int64_t)(i / (size_t)cd.y);
        const int64_t x = r % cd.x, z = r / cd.x;
        const int64_t y1 = min(2 * y + 1, fd.y - 1), x1 = min(2 * x + 1, fd.x - 1), z1 = min(2 * z + 1, fd.z - 1);
        const int64_t y0 = 2 * y, x0 = 2 * x, z0 = 2 * z;
        uint8_t m = fine[(z0 * fd.x + x0) * fd.y + y0];
        m = max(m, fine[(z0 * fd.x + x0) * fd.y + y1]);
        m = max(m, fine[(z0 * fd.x + x1) * fd.y + y0]);
        m = max(m, fine[(z0 * fd.x + x1) * fd.y + y1]);
        m = max(m, fine[(z1 * fd.x + x0) * fd.y + y0]);
        m = max(m, fine[(z1 * fd.x + x0) * fd.y + y1]);
        m = max(m, fine[(z1 * fd.x + x1) * fd.y + y0]);
        m = max(m, fine[(z1 * fd.x + x1) * fd.y + y1]);
        if (coarse) coarse[i] = m;
        bool seed = (mode == 0) ? (m >= level) : (mode == 1) ? (m <= level) : (m == level);
        pct[i] = seed ? SEED : EMPTY;
    }
}

/* seeds of the top level straight from the level map (PullingScheme::fill, k == l_max) */
__device__ __forceinline__ uint8_t seed_of(uint8_t m, int level, int mode) {
    const bool seed = (mode == 0) ? (m >= level) : (mode == 1) ? (m <= level) : (m == level);
    return seed ? SEED : EMPTY;
}
/* eight cells per thread (8-byte loads and stores) over the 8-aligned body, then the tail */
__device__ __forceinline__ void seed_body(const uint8_t* __restrict__ lev, uint8_t* pct, size_t n, int level, int mode, int64_t G,
                                          int64_t t) {
    const bool wide = (((uintptr_t)lev | (uintptr_t)pct) & 7u) == 0;
    const size_t nw = wide ? n / 8 : 0;
    const size_t stride = (size_t)G * blockDim.x;
    for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < nw; i += stride) {
        const uint64_t w = ((const uint64_t*)lev)[i];
        uint64_t o = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) o |= (uint64_t)seed_of((uint8_t)(w >> (8 * j)), level, mode) << (8 * j);
        ((uint64_t*)pct)[i] = o;
    }
    for (size_t i = nw * 8 + (size_t)t * blockDim.x + threadIdx.x; i < n; i += stride) pct[i] = seed_of(lev[i], level, mode);
}

/* one pulling-scheme level (PullingScheme.hpp:111-132 loop body): set_ascendant_neighbours
   and fill_neighbours on this level, set_filler on the finer one, fill_parent on the coarser
   one.  One lane per 8 consecutive y cells: the 3x3x3 neighbourhood
   predicates are formed as bit masks over the 9 neighbouring rows (10 bits: y0-1 .. y0+8),
   dilated along y by shifts, so a lane loads 9 row words instead of 27 bytes per cell. */
__device__ __forceinline__ void load_row8(const uint8_t* row, int yn, int y0, bool aligned8, uint8_t b[10]) {
    b[0] = (y0 > 0) ? row[y0 - 1] : (uint8_t)0;
    if (aligned8 && y0 + 8 <= yn) {
        const uint64_t w = *(const uint64_t*)(row + y0);
#pragma unroll
        for (int j = 0; j < 8; ++j) b[1 + j] = (uint8_t)(w >> (8 * j));
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) b[1 + j] = (y0 + j < yn) ? row[y0 + j] : (uint8_t)0;
    }
    b[9] = (y0 + 8 < yn) ? row[y0 + 8] : (uint8_t)0;
}

/* bit masks (bit k <-> y0 - 1 + k) of ASCENDANT and of SEED-or-PROPOGATE cells in a row */
__device__ __forceinline__ void row_masks(const uint8_t b[10], uint32_t& am, uint32_t& sm) {
    am = 0;
    sm = 0;
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        am |= (uint32_t)(b[k] == ASC) << k;
        sm |= (uint32_t)((b[k] == SEED) | (b[k] == PROP)) << k;
    }
}

/* One lane owns 8 consecutive y cells of one (x) column over a run of ZC planes and marches
   along z, keeping the ASCENDANT / SEED-or-PROPOGATE masks of the planes z-1, z, z+1 (each the
   OR over rows x-1, x, x+1): every plane's three rows are loaded once instead of three times.
   A cell's state is written only by its owner lane, so the cached bytes of the own row are
   current. */
constexpr int PULL_ZC = 4;   /* planes per lane of a single volume; a batch plan also times 8 */
template <int ZC>
__device__ __forceinline__ void pull_level8_body(uint8_t* P, Dims d, uint8_t* Pc, Dims dc, uint8_t* Pp, Dims dp, int is_top,
                                                 int has_parent, int64_t G, int64_t t) {
    const int yn = (int)d.y, xn = (int)d.x, zn = (int)d.z;
    const int nwr = (yn + 7) >> 3;
    const int nzc = (zn + ZC - 1) / ZC;
    const int64_t total = (int64_t)nwr * xn * nzc;
    const bool aligned8 = (yn & 7) == 0 && ((uintptr_t)P & 7) == 0;
    /* parent rows whose words never straddle two rows */
    const bool par_words = has_parent && (dp.y & 3) == 0 && ((uintptr_t)Pp & 3) == 0;
    for (int64_t i = t * blockDim.x + threadIdx.x; i < total; i += G * blockDim.x) {
        const int w = (int)(i % nwr);
        const int64_t r = i / nwr;
        const int x = (int)(r % xn), zc = (int)(r / xn);
        const int y0 = w * 8;
        const int z_begin = zc * ZC, z_end = min(zn, z_begin + ZC);
        /* masks of plane zz (0 outside the volume); own row bytes of that plane in own[] */
        auto plane = [&](int zz, uint32_t& am, uint32_t& sm, uint8_t own[10]) {
            am = 0;
            sm = 0;
            if (zz < 0 || zz >= zn) return;
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const int xx = x + dx;
                if (xx < 0 || xx >= xn) continue;
                uint8_t b[10];
                load_row8(P + ((int64_t)zz * xn + xx) * yn, yn, y0, aligned8, b);
                uint32_t a1, s1;
                row_masks(b, a1, s1);
                am |= a1;
                sm |= s1;
                if (dx == 0)
#pragma unroll
                    for (int k = 0; k < 10; ++k) own[k] = b[k];
            }
        };
        uint32_t pa0, ps0, pa1, ps1, pa2, ps2;
        uint8_t cur[10], nxt[10];
        plane(z_begin - 1, pa0, ps0, nxt);
        plane(z_begin, pa1, ps1, cur);
        for (int z = z_begin; z < z_end; ++z) {
            plane(z + 1, pa2, ps2, nxt);
            const uint32_t am = pa0 | pa1 | pa2, sm = ps0 | ps1 | ps2;
            const uint32_t ad = am | (am << 1) | (am >> 1), sd = sm | (sm << 1) | (sm >> 1);
            uint8_t* row = P + ((int64_t)z * xn + x) * yn;
            uint8_t nb[8];
            bool changed = false;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint8_t old = cur[1 + j];
                const bool any_asc = (ad >> (j + 1)) & 1u, any_sp = (sd >> (j + 1)) & 1u;
                uint8_t nw = old;
                if (!is_top) {
                    if (old == EMPTY) nw = any_asc ? ASCN : (any_sp ? BOUNDARY : EMPTY);
                    else if (old == SEED) nw = any_asc ? PROP : SEED;
                } else if (old == EMPTY && any_sp) {
                    nw = BOUNDARY;
                }
                nb[j] = nw;
                changed |= (nw != old) && (y0 + j < yn);
            }
            if (changed) {
                if (aligned8) {
                    uint64_t v = 0;
#pragma unroll
                    for (int j = 0; j < 8; ++j) v |= (uint64_t)nb[j] << (8 * j);
                    *(uint64_t*)(row + y0) = v;
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        if (y0 + j < yn && nb[j] != cur[1 + j]) row[y0 + j] = nb[j];
                }
            }
            /* fill_parent for the lane's 4 parent cells with one word: the parent level holds only
               EMPTY, SEED and ASCENDANT now (its seeds, and ASCENDANT from this pass), so OR-ing
               ASCENDANT's bits into the non-SEED bytes wanted gives each the reference's value, in
               any order between lanes that share the word */
            const bool word_parent = has_parent && par_words;
            if (word_parent) {
                uint32_t want = 0;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint8_t nw = nb[j];
                    if (y0 + j < yn && (nw == SEED || nw == PROP || nw == ASC)) want |= 0xFFu << (8 * (j >> 1));
                }
                if (want) {
                    uint32_t* pw = (uint32_t*)&Pp[((int64_t)(z / 2) * dp.x + (x / 2)) * dp.y + (y0 / 2)];
                    const uint32_t cur4 = *pw;
                    uint32_t mask = 0;
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (((want >> (8 * q)) & 0xFFu) && ((cur4 >> (8 * q)) & 0xFFu) != SEED) mask |= (uint32_t)ASC << (8 * q);
                    if (mask) atomicOr(pw, mask);
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int y = y0 + j;
                if (y >= yn) break;
                const uint8_t nw = nb[j];
                /* set_filler: empty children of ASCENDANTNEIGHBOUR / PROPOGATE become FILLER */
                if (!is_top && (nw == ASCN || nw == PROP)) {
                    const int cz = (z == zn - 1 && (dc.z & 1)) ? 1 : 2;
                    const int cx = (x == xn - 1 && (dc.x & 1)) ? 1 : 2;
                    const int cy = (y == yn - 1 && (dc.y & 1)) ? 1 : 2;
                    for (int a = 0; a < cz; ++a)
                        for (int bb = 0; bb < cx; ++bb)
                            for (int c = 0; c < cy; ++c) {
                                uint8_t* ch = &Pc[((int64_t)(2 * z + a) * dc.x + (2 * x + bb)) * dc.y + (2 * y + c)];
                                if (*ch == EMPTY) *ch = FILLER;
                            }
                }
                /* fill_parent */
                if (!word_parent && has_parent && (nw == SEED || nw == PROP || nw == ASC)) {
                    uint8_t* pp = &Pp[((int64_t)(z / 2) * dp.x + (x / 2)) * dp.y + (y / 2)];
                    if (*pp != SEED) *pp = ASC;
                }
            }
            pa0 = pa1;
            ps0 = ps1;
            pa1 = pa2;
            ps1 = ps2;
#pragma unroll
            for (int k = 0; k < 10; ++k) cur[k] = nxt[k];
        }
    }
}

/* equivalence optimisation step for one level (LinearAccess.hpp:265-294) */
__device__ __forceinline__ void upseed_body(uint8_t* P, Dims d, const uint8_t* __restrict__ Pp, Dims dp, int min_type, int64_t G,
                                            int64_t t, const FastDiv& fy, const FastDiv& fx, bool i32) {
    const size_t n = (size_t)(d.y * d.x * d.z);
    for (size_t i = (size_t)t * blockDim.x + threadIdx.x; i < n; i += (size_t)G * blockDim.x) {
        int64_t y, x, z;
        if (i32) {
            const uint32_t r = fy.div((uint32_t)i), zz = fx.div(r);
            y = (int64_t)((uint32_t)i - r * (uint32_t)d.y);
            x = (int64_t)(r - zz * (uint32_t)d.x);
            z = zz;
        } else {
            y = (int64_t)(i % (size_t)d.y);
            const int64_t r = (int64_t)(i / (size_t)d.y);
            x = r % d.x;
            z = r / d.x;
        }
        const uint8_t s = Pp[((z / 2) * dp.x + (x / 2)) * dp.y + (y / 2)];
        if (s > 0 && s <= min_type) P[i] = UPSEED;
    }
}

/* row table of the APR levels l_min..l_max and the particle-cell maps that define them */
constexpr int MAXL = 24;
struct RowTable {
    int n_levels;                 /* APR levels l_min..l_max           */
    int l_min, l_max;
    int64_t row_begin[MAXL];      /* xz_end index of row 0 of level    */
    int64_t rows_x[MAXL];         /* x_num of the APR level            */
    int64_t map_off[MAXL];        /* byte offset of the pct map used   */
    int64_t map_y[MAXL], map_x[MAXL]; /* dims of that pct map           */
    int64_t y_max;                /* y_num at l_max                    */
    int64_t row_prefix[MAXL + 1]; /* first global row of each APR level */
    int min_type;
    /* the l_max rows (z, x) read pct row (z/2, x/2) of l_max - 1, so up to four of them hold
       identical particles: work items are the rows below l_max, then one per such pct row */
    int64_t top_z;                /* z_num at l_max                    */
    int64_t top_pz;               /* z_num at l_max - 1                */
    int64_t n_work;               /* row_prefix[top] + map_x * top_pz  */
};

/* weight of one pct cell for APR level li (0 or 1; 0..2 at l_max) */
__device__ __forceinline__ int cell_weight(const RowTable& t, int li, uint8_t st, int64_t yy) {
    if (li == t.n_levels - 1) {
        if (st > 0 && st <= t.min_type) return 1 + ((2 * yy < t.y_max - 1) ? 1 : 0);
        return 0;
    }
    return (st > t.min_type && st <= UPSEED) ? 1 : 0;
}

__device__ __forceinline__ int row_level_of(const RowTable& t, int64_t r) {
    int li = 0;
    while (li + 1 < t.n_levels && r >= t.row_prefix[li + 1]) ++li;
    return li;
}

constexpr int CPL = 16;   /* cells per lane per chunk: a wave covers 1024 cells of a row */

/* weights of the CPL cells [c0, c0 + CPL) of a row (0 beyond its end) */
__device__ __forceinline__ void chunk_weights(const RowTable& t, int li, const uint8_t* p, int ylen, int c0, int w[CPL]) {
    const bool top = (li == t.n_levels - 1);
    if (c0 + CPL <= ylen && ((uintptr_t)(p + c0) & 15) == 0) {
        const uint4 v = *(const uint4*)(p + c0);
        const uint32_t words[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < CPL; ++j) w[j] = cell_weight(t, li, (uint8_t)(words[j >> 2] >> (8 * (j & 3))), c0 + j);
    } else {
#pragma unroll
        for (int j = 0; j < CPL; ++j) w[j] = (c0 + j < ylen) ? cell_weight(t, li, p[c0 + j], c0 + j) : 0;
    }
    (void)top;
}

/* the APR rows of work item w (1, or up to 4 sharing a pct row at l_max) and its cells */
struct WorkRows {
    int li;
    const uint8_t* p;     /* pct row          */
    int ylen;
    int n;                /* destination rows */
    int64_t r[4];
};
__device__ __forceinline__ WorkRows work_rows(const RowTable& t, const uint8_t* maps, int64_t w) {
    WorkRows o;
    const int top = t.n_levels - 1;
    if (w < t.row_prefix[top]) {
        o.li = row_level_of(t, w);
        o.p = maps + t.map_off[o.li] + (w - t.row_prefix[o.li]) * t.map_y[o.li];
        o.n = 1;
        o.r[0] = w;
    } else {
        const int64_t prow = w - t.row_prefix[top];
        const int64_t zm = prow / t.map_x[top], xm = prow - zm * t.map_x[top];
        o.li = top;
        o.p = maps + t.map_off[top] + prow * t.map_y[top];
        /* rows (2zm, 2xm), (2zm, 2xm+1), (2zm+1, 2xm), (2zm+1, 2xm+1) that exist, in that order */
        const bool hx = 2 * xm + 1 < t.rows_x[top], hz = 2 * zm + 1 < t.top_z;
        const int64_t r00 = t.row_prefix[top] + 2 * zm * t.rows_x[top] + 2 * xm;
        o.r[0] = r00;
        o.r[1] = hx ? r00 + 1 : r00 + t.rows_x[top];
        o.r[2] = hx ? r00 + t.rows_x[top] : r00;
        o.r[3] = r00 + t.rows_x[top] + 1;
        o.n = 1 + (hx ? 1 : 0) + (hz ? 1 : 0) + ((hx && hz) ? 1 : 0);
    }
    o.ylen = (int)t.map_y[o.li];
    return o;
}

/* particles of each APR row: 16 lanes per work item, 256 cells per step (short rows of small
   volumes and coarse levels keep the lanes busy; long rows loop) */
constexpr int COUNT_LANES = 16;
__device__ __forceinline__ void row_count_body(const uint8_t* __restrict__ maps, const RowTable& t, int64_t nrows, uint32_t* cnt,
                                               int64_t G, int64_t blk) {
    const int lane = threadIdx.x & (COUNT_LANES - 1);
    const int64_t wid = (blk * blockDim.x + threadIdx.x) / COUNT_LANES;
    const int64_t nw = (G * blockDim.x) / COUNT_LANES;
    (void)nrows;
    for (int64_t w = wid; w < t.n_work; w += nw) {
        const WorkRows wr = work_rows(t, maps, w);
        uint32_t c = 0;
        for (int base = 0; base < wr.ylen; base += COUNT_LANES * CPL) {
            int wt[CPL];
            chunk_weights(t, wr.li, wr.p, wr.ylen, base + lane * CPL, wt);
#pragma unroll
            for (int j = 0; j < CPL; ++j) c += (uint32_t)wt[j];
        }
        /* the 16 lanes of a group share w, so they are active together */
        for (int o = COUNT_LANES / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, COUNT_LANES);
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < wr.n) cnt[wr.r[k]] = c;
        }
    }
}

/* y values of each APR row from its scanned offset.  A group of FILL_LANES lanes (a whole wave
   for long rows, 16 lanes for short ones) takes one work item; a lane takes CPL consecutive
   cells of a FILL_LANES * CPL-cell chunk; a prefix sum over the group places each lane's values
   in the group's LDS stage, which the group then copies to y_vec with consecutive lanes on
   consecutive entries (once per destination row). */
constexpr int FILL_WAVES = 4;
/* level images for a row fill that also samples: APR level index li -> image of that level
   (the converted image at l_max, the mean pyramid below), rows in (z, x) order like the APR rows */
struct RowImages {
    const uint16_t* img[MAXL];
    int64_t y[MAXL];
    uint16_t* parts;   /* nullptr: no sampling */
};

template <int FILL_LANES>
__device__ __forceinline__ void row_fill_body(const uint8_t* __restrict__ maps, const RowTable& t, int64_t nrows,
                                              const uint64_t* __restrict__ row_off, uint16_t* y_vec, const RowImages& ri, int64_t G,
                                              int64_t blk) {
    constexpr int FILL_GROUPS = FILL_WAVES * 64 / FILL_LANES;
    __shared__ uint16_t stage[FILL_GROUPS][2 * FILL_LANES * CPL];
    const int lane = threadIdx.x & (FILL_LANES - 1);
    uint16_t* st = stage[threadIdx.x / FILL_LANES];
    const int64_t wid = (blk * blockDim.x + threadIdx.x) / FILL_LANES;
    const int64_t nw = (G * blockDim.x) / FILL_LANES;
    (void)nrows;
    /* the 16 lanes of a group share w (and every loop bound below), so they are active together */
    for (int64_t w = wid; w < t.n_work; w += nw) {
        const WorkRows wr = work_rows(t, maps, w);
        const bool top = (wr.li == t.n_levels - 1);
        uint64_t pos[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < wr.n) pos[k] = row_off[wr.r[k]];
        for (int base = 0; base < wr.ylen; base += FILL_LANES * CPL) {
            int wt[CPL];
            const int c0 = base + lane * CPL;
            chunk_weights(t, wr.li, wr.p, wr.ylen, c0, wt);
            int c = 0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) c += wt[j];
            /* inclusive prefix over the group */
            int incl = c;
            for (int o = 1; o < FILL_LANES; o <<= 1) {
                const int v = __shfl_up(incl, o, FILL_LANES);
                if (lane >= o) incl += v;
            }
            const int total = __shfl(incl, FILL_LANES - 1, FILL_LANES);
            int q = incl - c;
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                const int yy = c0 + j;
                if (wt[j] >= 1) st[q++] = (uint16_t)(top ? 2 * yy : yy);
                if (wt[j] == 2) st[q++] = (uint16_t)(2 * yy + 1);
            }
            __threadfence_block();
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (k >= wr.n) break;
                uint16_t* dst = y_vec + pos[k];
                if (ri.parts) {
                    /* sample_image's gather (ParticleData.hpp:221-261) for this row's particles */
                    const uint16_t* irow = ri.img[wr.li] + (wr.r[k] - t.row_prefix[wr.li]) * ri.y[wr.li];
                    uint16_t* pdst = ri.parts + pos[k];
                    int i = lane;
                    /* four particles per lane in flight */
                    for (; i + 3 * FILL_LANES < total; i += 4 * FILL_LANES) {
                        const uint16_t y0 = st[i], y1 = st[i + FILL_LANES], y2 = st[i + 2 * FILL_LANES], y3 = st[i + 3 * FILL_LANES];
                        const uint16_t v0 = irow[y0], v1 = irow[y1], v2 = irow[y2], v3 = irow[y3];
                        dst[i] = y0;
                        dst[i + FILL_LANES] = y1;
                        dst[i + 2 * FILL_LANES] = y2;
                        dst[i + 3 * FILL_LANES] = y3;
                        pdst[i] = v0;
                        pdst[i + FILL_LANES] = v1;
                        pdst[i + 2 * FILL_LANES] = v2;
                        pdst[i + 3 * FILL_LANES] = v3;
                    }
                    for (; i < total; i += FILL_LANES) {
                        const uint16_t yv = st[i];
                        dst[i] = yv;
                        pdst[i] = irow[yv];
                    }
                } else {
                    for (int i = lane; i < total; i += FILL_LANES) dst[i] = st[i];
                }
                pos[k] += (uint64_t)total;
            }
            __threadfence_block();   /* the stage is rewritten by the next chunk */
        }
    }
}

/* xz_end_vec of each row: its offset + its count */
__device__ __forceinline__ void rows_xz_body(const uint64_t* __restrict__ row_off, const uint32_t* __restrict__ cnt, const RowTable& t,
                                             uint64_t* xz_end, int64_t nrows, int64_t blk) {
    const int64_t r = blk * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int li = row_level_of(t, r);
    xz_end[t.row_begin[li] + (r - t.row_prefix[li])] = row_off[r] + cnt[r];
}

/* exclusive scan, three phases over blocks of 1024 */
__device__ __forceinline__ void scan_blocks_body(const uint32_t* __restrict__ in, uint64_t* out, uint64_t* block_sums, int64_t n,
                                                 int64_t blk) {
    __shared__ uint64_t sh[1024];
    const int64_t i = blk * 1024 + threadIdx.x;
    uint64_t v = i < n ? in[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        uint64_t a = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += a;
        __syncthreads();
    }
    if (i < n) out[i] = sh[threadIdx.x] - v;
    if (threadIdx.x == 1023) block_sums[blk] = sh[1023];
    __syncthreads();   /* sh is reused by the next tile of a persistent block */
}

__device__ __forceinline__ void scan_sums_body(uint64_t* block_sums, int64_t nb, uint64_t* total) {
    /* single block: exclusive scan of the block sums (serial chunks per thread) */
    __shared__ uint64_t sh[1024];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b = (int64_t)threadIdx.x * per;
    uint64_t s = 0;
    for (int64_t k = b; k < min(b + per, nb); ++k) s += block_sums[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        uint64_t a = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += a;
        __syncthreads();
    }
    uint64_t run = sh[threadIdx.x] - s;
    for (int64_t k = b; k < min(b + per, nb); ++k) {
        uint64_t v = block_sums[k];
        block_sums[k] = run;
        run += v;
    }
    if (threadIdx.x == 1023) *total = sh[1023];
}

__device__ __forceinline__ void scan_add_body(uint64_t* out, const uint64_t* __restrict__ block_sums, int64_t n, int64_t blk) {
    const int64_t i = blk * blockDim.x + threadIdx.x;
    if (i < n) out[i] += block_sums[i / 1024];
}

/* tile bodies of the pulling scheme and the structure */
struct Seed {
    struct Args {
        const uint8_t* lev;
        uint8_t* pct;
        size_t n;
        int level, mode;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "pulling_seed"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { seed_body(a.lev, a.pct, a.n, a.level, a.mode, a.G, t); }
};
struct LevelDown {
    struct Args {
        const uint8_t* fine;
        Dims fd;
        uint8_t* coarse;
        Dims cd;
        uint8_t* pct;
        int level, mode;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "pulling_level_down"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { level_down_body(a.fine, a.fd, a.coarse, a.cd, a.pct, a.level, a.mode, a.G, t); }
};
struct PullLevelArgs {
    uint8_t* P;
    Dims d;
    uint8_t* Pc;
    Dims dc;
    uint8_t* Pp;
    Dims dp;
    int is_top, has_parent;
    int64_t G;
};
template <int ZC = PULL_ZC>
struct PullLevel {
    using Args = PullLevelArgs;
    static constexpr int THREADS = 256;
    static const char* name() { return "pulling_level"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        pull_level8_body<ZC>(a.P, a.d, a.Pc, a.dc, a.Pp, a.dp, a.is_top, a.has_parent, a.G, t);
    }
};
struct Upseed {
    struct Args {
        uint8_t* P;
        Dims d;
        const uint8_t* Pp;
        Dims dp;
        int min_type;
        int64_t G;
        FastDiv fy, fx;
        int i32;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "upseed"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        upseed_body(a.P, a.d, a.Pp, a.dp, a.min_type, a.G, t, a.fy, a.fx, a.i32);
    }
};
struct RowCount {
    struct Args {
        const uint8_t* maps;
        RowTable t;
        int64_t nrows;
        uint32_t* cnt;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "row_count"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { row_count_body(a.maps, a.t, a.nrows, a.cnt, a.G, t); }
};
template <int LANES>
struct RowFill {
    struct Args {
        const uint8_t* maps;
        RowTable t;
        int64_t nrows;
        const uint64_t* row_off;
        uint16_t* y_vec;
        int64_t G;
        RowImages ri;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return LANES == 64 ? "row_fill" : "row_fill_short"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        row_fill_body<LANES>(a.maps, a.t, a.nrows, a.row_off, a.y_vec, a.ri, a.G, t);
    }
};
struct RowsXz {
    struct Args {
        const uint64_t* row_off;
        const uint32_t* cnt;
        RowTable t;
        uint64_t* xz_end;
        int64_t nrows;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "rows_xz"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { rows_xz_body(a.row_off, a.cnt, a.t, a.xz_end, a.nrows, t); }
};
struct ScanBlocks {
    struct Args {
        const uint32_t* in;
        uint64_t* out;
        uint64_t* block_sums;
        int64_t n;
    };
    static constexpr int THREADS = 1024;
    static const char* name() { return "scan_blocks"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { scan_blocks_body(a.in, a.out, a.block_sums, a.n, t); }
};
struct ScanSums {
    struct Args {
        uint64_t* block_sums;
        int64_t nb;
        uint64_t* total;
    };
    static constexpr int THREADS = 1024;
    static const char* name() { return "scan_sums"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t) { scan_sums_body(a.block_sums, a.nb, a.total); }
};
struct ScanAdd {
    struct Args {
        uint64_t* out;
        const uint64_t* block_sums;
        int64_t n;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "scan_add"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) { scan_add_body(a.out, a.block_sums, a.n, t); }
};

/* mean of the 2x2x2 block in the reference's summation order (PixelData.hpp:952-962), stored
   as the pyramid's element type (u16: truncation like a float -> uint16 assignment) */
template <typename T> __device__ __forceinline__ T pyr_store(float m);
template <> __device__ __forceinline__ uint16_t pyr_store<uint16_t>(float m) { return (uint16_t)(int32_t)m; }
template <> __device__ __forceinline__ float pyr_store<float>(float m) { return m; }

template <typename T>
__global__ void k_pyr_down(const T* __restrict__ fine, Dims fd, T* coarse, Dims cd) {
    const size_t n = (size_t)(cd.y * cd.x * cd.z);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int64_t y = (int64_t)(i % (size_t)cd.y);
        const int64_t r = (int64_t)(i / (size_t)cd.y);
        const int64_t x = r % cd.x, z = r / cd.x;
        const int64_t y1 = min(2 * y + 1, fd.y - 1), x1 = min(2 * x + 1, fd.x - 1), z1 = min(2 * z + 1, fd.z - 1);
        const int64_t y0 = 2 * y, x0 = 2 * x, z0 = 2 * z;
        float s = (float)fine[(z0 * fd.x + x0) * fd.y + y0];
        s = s + (float)fine[(z0 * fd.x + x0) * fd.y + y1];
        s = s + (float)fine[(z0 * fd.x + x1) * fd.y + y0];
        s = s + (float)fine[(z0 * fd.x + x1) * fd.y + y1];
        s = s + (float)fine[(z1 * fd.x + x0) * fd.y + y0];
        s = s + (float)fine[(z1 * fd.x + x0) * fd.y + y1];
        s = s + (float)fine[(z1 * fd.x + x1) * fd.y + y0];
        s = s + (float)fine[(z1 * fd.x + x1) * fd.y + y1];
        const float m = s / 8.0f;
        coarse[i] = pyr_store<T>(m);
    }
}

/* the same mean for u16 with 8 fine y per lane (16-byte row loads, 4 coarse outputs); the
   integer sum of eight u16 values is exact in float, so the order of the float additions
   above does not matter and the result is identical */
__global__ __launch_bounds__(256) void k_pyr_down_u16v(const uint16_t* __restrict__ fine, Dims fd, uint16_t* coarse, Dims cd) {
    const int64_t ng = cd.y / 4;
    const int64_t total = ng * cd.x * cd.z;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i % ng, r = i / ng;
        const int64_t x = r % cd.x, z = r / cd.x;
        const int64_t x0 = 2 * x, x1 = min(2 * x + 1, fd.x - 1), z0 = 2 * z, z1 = min(2 * z + 1, fd.z - 1);
        const uint4 a = *(const uint4*)(fine + (z0 * fd.x + x0) * fd.y + 8 * g);
        const uint4 b = *(const uint4*)(fine + (z0 * fd.x + x1) * fd.y + 8 * g);
        const uint4 c = *(const uint4*)(fine + (z1 * fd.x + x0) * fd.y + 8 * g);
        const uint4 d = *(const uint4*)(fine + (z1 * fd.x + x1) * fd.y + 8 * g);
        const uint32_t wa[4] = {a.x, a.y, a.z, a.w}, wb[4] = {b.x, b.y, b.z, b.w}, wc[4] = {c.x, c.y, c.z, c.w},
                       wd[4] = {d.x, d.y, d.z, d.w};
        uint32_t o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t sum = (wa[j] & 0xFFFFu) + (wa[j] >> 16) + (wb[j] & 0xFFFFu) + (wb[j] >> 16) + (wc[j] & 0xFFFFu) +
                                 (wc[j] >> 16) + (wd[j] & 0xFFFFu) + (wd[j] >> 16);
            o[j] = (uint32_t)pyr_store<uint16_t>((float)sum / 8.0f);
        }
        *(uint2*)(coarse + (z * cd.x + x) * cd.y + 4 * g) = make_uint2(o[0] | (o[1] << 16), o[2] | (o[3] << 16));
    }
}

template <typename T>
struct LevelPtrs {
    const T* img[MAXL];
    int64_t y[MAXL], x[MAXL];
    int64_t row_begin[MAXL];
    int64_t nrows[MAXL];
    int64_t rows_prefix[MAXL + 1];
    int n_levels;
};

/* one wave per row; lanes stride the row's particles */
template <typename T>
__global__ __launch_bounds__(256) void k_gather(LevelPtrs<T> L, const uint64_t* __restrict__ xz_end, const uint16_t* __restrict__ y_vec,
                                                T* parts, int64_t total_rows) {
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / 64;
    for (int64_t r = wid; r < total_rows; r += nw) {
        int li = 0;
        while (li + 1 < L.n_levels && r >= L.rows_prefix[li + 1]) ++li;
        const int64_t row = r - L.rows_prefix[li];
        const int64_t xzi = L.row_begin[li] + row;
        const uint64_t b = xz_end[xzi - 1], e = xz_end[xzi];
        const int64_t z = row / L.x[li], x = row % L.x[li];
        const T* plane = L.img[li] + (z * L.x[li] + x) * L.y[li];
        /* four independent particles per lane in flight */
        uint64_t i = b + lane;
        for (; i + 192 < e; i += 256) {
            const uint16_t y0 = y_vec[i], y1 = y_vec[i + 64], y2 = y_vec[i + 128], y3 = y_vec[i + 192];
            const T v0 = plane[y0], v1 = plane[y1], v2 = plane[y2], v3 = plane[y3];
            parts[i] = v0;
            parts[i + 64] = v1;
            parts[i + 128] = v2;
            parts[i + 192] = v3;
        }
        for (; i < e; i += 64) parts[i] = plane[y_vec[i]];
    }
}

}  // namespace

void exclusive_scan_u32_async(aprh_context* ctx, const uint32_t* in, uint64_t* out, int64_t n, uint64_t* d_total) {
    if (n <= 0) {
        launch_zero(ctx, d_total, 8);
        return;
    }
    const int64_t nb = cdiv(n, 1024);
    Buf bsum(ctx->pool, (size_t)(nb + 1) * 8);
    launch_tiles<ScanBlocks>(ctx, {in, out, bsum.as<uint64_t>(), n}, nb);
    launch_tiles<ScanSums>(ctx, {bsum.as<uint64_t>(), nb, d_total}, 1);
    launch_tiles<ScanAdd>(ctx, {out, bsum.as<uint64_t>(), n}, cdiv(n, 256));
}

uint64_t exclusive_scan_u32(aprh_context* ctx, const uint32_t* in, uint64_t* out, int64_t n) {
    if (n <= 0) return 0;
    hipStream_t st = ctx->stream;
    uint64_t* d_total = (uint64_t*)ctx->d_flags + 8;
    exclusive_scan_u32_async(ctx, in, out, n, d_total);
    uint64_t total = 0;
    HIP_CHECK(hipMemcpyAsync(&total, d_total, 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return total;
}

template <typename T>
std::vector<int64_t> build_pyramid(aprh_context* ctx, const aprh_apr& a, const T* img, Buf& pyr);
template <typename T>
void pyramid_levels(aprh_context* ctx, const aprh_apr& a, const T* img, T* Pyr, const std::vector<int64_t>& off, int first);

/* layout of the particle-cell tree levels lmin .. lmax-1 in one u8 buffer (8-byte aligned
   levels, so whole-word passes apply): element offsets, off[top + 1] = total */
std::vector<int64_t> pct_layout(const aprh_apr& a) {
    const int top = a.l_max - 1;
    std::vector<int64_t> off(std::max(top + 2, 1), 0);
    for (int l = a.l_min; l <= top; ++l) off[l + 1] = off[l] + (a.y_num[l] * a.x_num[l] * a.z_num[l] + 7) / 8 * 8;
    return off;
}

/* PullingScheme::fill of every level (LocalParticleCellSet.hpp:58-91): seeds of the top level
   straight from the level map, then the 2x2x2 max pyramid of the map with the seeds of each
   coarser level.  zn: z extent per level (the APR's, or a patch's) */
void seed_levels(aprh_context* ctx, const aprh_apr& a, const uint8_t* level_top, const std::vector<int64_t>& zn, uint8_t* P,
                 const std::vector<int64_t>& off) {
    const int top = a.l_max - 1, lmin = a.l_min;
    auto D = [&](int l) { return Dims{a.y_num[l], a.x_num[l], zn[l]}; };
    auto mode_of = [&](int l) { return l == top ? 0 : (l == lmin ? 1 : 2); };
    Buf lev(ctx->pool, (size_t)std::max<int64_t>(off[top + 1], 8));
    uint8_t* LV = lev.as<uint8_t>();
    {
        const size_t n = (size_t)(a.y_num[top] * a.x_num[top] * zn[top]);
        const int64_t G = stride_tiles((int64_t)n, 256 * 4);
        launch_tiles<Seed>(ctx, {level_top, P + off[top], n, top, mode_of(top), G}, G);
    }
    const uint8_t* fine = level_top;
    for (int l = top - 1; l >= lmin; --l) {
        const int64_t G = stride_tiles(off[l + 1] - off[l], 256 * 4);
        launch_tiles<LevelDown>(ctx, {fine, D(l + 1), LV + off[l], D(l), P + off[l], l, mode_of(l), G}, G);
        fine = LV + off[l];
    }
}

namespace {
/* seeds of one patch's plane range into the global particle-cell tree level:
   global(z + oz) |= SEED where local(z) == SEED, z in [z_lo, z_hi), z + oz < global z extent */
struct SeedMerge {
    struct Args {
        const uint8_t* local;
        uint8_t* global;
        int64_t plane;        /* y * x of the level */
        int64_t z_lo, z_hi, oz, gz;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "seed_merge"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        const int64_t zh = min(a.z_hi, a.gz - a.oz);
        if (zh <= a.z_lo) return;
        const int64_t n = (zh - a.z_lo) * a.plane;
        const uint8_t* src = a.local + a.z_lo * a.plane;
        uint8_t* dst = a.global + (a.z_lo + a.oz) * a.plane;
        for (int64_t i = t * blockDim.x + threadIdx.x; i < n; i += a.G * blockDim.x)
            if (src[i] == SEED) dst[i] = SEED;
    }
};

/* blocked sampling of one level: rows (z + oz, x) for patch planes z in [z_lo, z_hi) take the
   patch pyramid's values at (y, x, z) (ParticleData.hpp:329-362); one wave per row */
template <typename T>
struct GatherPatch {
    struct Args {
        const T* img;                 /* patch pyramid level, (z, x, y) with y fastest */
        int64_t Y, X;
        const uint64_t* xz_end;
        const uint16_t* y_vec;
        int64_t row_begin;            /* xz_end index of the level's row (0, 0) */
        int64_t z_lo, z_hi, oz;
        T* parts;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "gather_patch"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        const int lane = threadIdx.x & 63;
        const int64_t nrows = (a.z_hi - a.z_lo) * a.X;
        const int64_t wpb = blockDim.x / 64;
        for (int64_t r = t * wpb + (threadIdx.x >> 6); r < nrows; r += a.G * wpb) {
            const int64_t z = a.z_lo + r / a.X, x = r % a.X;
            const int64_t rid = a.row_begin + (z + a.oz) * a.X + x;
            const uint64_t b = a.xz_end[rid - 1], e = a.xz_end[rid];
            const T* row = a.img + (z * a.X + x) * a.Y;
            for (uint64_t i = b + lane; i < e; i += 64) {
                const int64_t y = a.y_vec[i];
                if (y < a.Y) a.parts[i] = row[y];
            }
        }
    }
};
}  // namespace

/* z extents per level of a patch of zp planes (the pyramid / particle-cell tree geometry of the
   patch: ceil(zp / 2^(l_max - l))) */
std::vector<int64_t> patch_z_extents(const aprh_apr& a, int64_t zp) {
    std::vector<int64_t> zn(a.l_max + 1, 0);
    for (int l = 0; l <= a.l_max; ++l) zn[l] = cdiv(zp, (int64_t)1 << (a.l_max - l));
    return zn;
}

/* PullingScheme::fill_patch for every level (APRConverterBatch.hpp:287-322,
   PullingScheme.hpp:175-244): the patch's own seeds (its level map and the max pyramid of it),
   merged into the global tree P over the patch's non-ghost planes of each level */
void seed_patch(aprh_context* ctx, const aprh_apr& a, uint8_t* P, const std::vector<int64_t>& off, const uint8_t* level_top_patch,
                int64_t zp, int64_t z0, int64_t gl, int64_t gr) {
    const int top = a.l_max - 1;
    const std::vector<int64_t> zn = patch_z_extents(a, zp);
    std::vector<int64_t> loff(top + 2, 0);
    for (int l = a.l_min; l <= top; ++l) loff[l + 1] = loff[l] + (a.y_num[l] * a.x_num[l] * zn[l] + 7) / 8 * 8;
    Buf local(ctx->pool, (size_t)std::max<int64_t>(loff[top + 1], 8));
    seed_levels(ctx, a, level_top_patch, zn, local.as<uint8_t>(), loff);
    for (int l = a.l_min; l <= top; ++l) {
        const int64_t f = (int64_t)1 << (a.l_max - l);
        const int64_t zl = gl / f, zr = gr / f, oz = z0 / f - zl;
        const int64_t plane = a.y_num[l] * a.x_num[l];
        const int64_t G = stride_tiles(std::max<int64_t>((zn[l] - zl - zr) * plane, 1), 256 * 4);
        launch_tiles<SeedMerge>(ctx, {local.as<uint8_t>() + loff[l], P + off[l], plane, zl, zn[l] - zr, oz, a.z_num[l], G}, G);
    }
}

/* ParticleData::sample_parts_from_img_blocked (ParticleData.hpp:292-362) on a device image:
   per z block the mean pyramid of the patch (ghosts included) and the rows of its non-ghost
   planes of each level; rows past a level's z extent are skipped; particles no block reaches
   stay 0 */
void sample_blocked_u16(aprh_context* ctx, const aprh_apr& a, const uint16_t* img, int64_t z_block_size, int64_t ghost_z,
                        uint16_t* parts) {
    launch_zero(ctx, parts, std::max<uint64_t>(a.total, 1) * 2);
    const int64_t Z = a.org[2], plane = a.org[0] * a.org[1];
    const int64_t nb = std::max<int64_t>(Z / z_block_size, 1);
    for (int64_t bk = 0; bk < nb; ++bk) {
        const int64_t z0 = bk * z_block_size, zf = (bk == nb - 1) ? Z : (bk + 1) * z_block_size;
        const int64_t gl = std::min(z0, ghost_z), gr = std::min(Z - zf, ghost_z);
        const int64_t zp = zf + gr - (z0 - gl);
        const uint16_t* pimg = img + (z0 - gl) * plane;
        aprh_apr pa;
        pa.l_min = a.l_min;
        pa.l_max = a.l_max;
        pa.y_num = a.y_num;
        pa.x_num = a.x_num;
        pa.z_num = patch_z_extents(a, zp);
        int64_t tot = 0;
        const std::vector<int64_t> poff = pyramid_layout(pa, &tot);
        Buf pyr(ctx->pool, (size_t)std::max<int64_t>(tot, 1) * 2);
        if (a.l_max > a.l_min) pyramid_levels<uint16_t>(ctx, pa, pimg, pyr.as<uint16_t>(), poff, a.l_max - 1);
        for (int l = a.l_min; l <= a.l_max; ++l) {
            const int64_t f = (int64_t)1 << (a.l_max - l);
            const int64_t zl = gl / f, zr = gr / f, oz = z0 / f - zl;
            const int64_t zh = std::min(pa.z_num[l] - zr, a.z_num[l] - oz);
            if (zh <= zl) continue;
            const uint16_t* lvimg = (l == a.l_max) ? pimg : pyr.as<uint16_t>() + poff[l];
            const int64_t G = stride_tiles((zh - zl) * a.x_num[l], 4, 16384);
            launch_tiles<GatherPatch<uint16_t>>(ctx, {lvimg, a.y_num[l], a.x_num[l], a.xz_end.as<uint64_t>(), a.y_vec.as<uint16_t>(),
                                                      (int64_t)a.level_xz_vec[l], zl, zh, oz, parts, G},
                                                G);
        }
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
}

void build_structure(aprh_context* ctx, aprh_apr& a, const uint8_t* level_top, const aprh_params& par) {
    build_structure_seeded(ctx, a, level_top, nullptr, par);
}

void build_structure_seeded(aprh_context* ctx, aprh_apr& a, const uint8_t* level_top, uint8_t* seeded, const aprh_params& par) {
    hipStream_t st = ctx->stream;
    const int lmax = a.l_max, lmin = a.l_min;
    if (lmax <= 2) {
        if (ctx->async_mode) throw Error(APRH_ERR_UNSUPPORTED, "batched conversion of volumes with l_max <= 2");
        /* LinearAccess.hpp:229-260: every pixel is a particle at l_max */
        const int64_t yL = a.y_num[lmax], xL = a.x_num[lmax], zL = a.z_num[lmax];
        std::vector<uint64_t> xz(a.n_xz, 0);
        uint64_t c = 0;
        for (int64_t z = 0; z < zL; ++z)
            for (int64_t x = 0; x < xL; ++x) {
                c += (uint64_t)yL;
                xz[a.level_xz_vec[lmax] + z * xL + x] = c;
            }
        /* the reference leaves the lower rows at 0 and skips the partial sum */
        std::vector<uint16_t> yv(c);
        uint64_t k = 0;
        for (int64_t z = 0; z < zL; ++z)
            for (int64_t x = 0; x < xL; ++x)
                for (int64_t y = 0; y < yL; ++y) yv[k++] = (uint16_t)y;
        a.total = c;
        a.xz_end.reset(ctx->pool, a.n_xz * 8);
        a.y_vec.reset(ctx->pool, std::max<uint64_t>(c, 1) * 2);
        HIP_CHECK(hipMemcpyAsync(a.xz_end.ptr(), xz.data(), a.n_xz * 8, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(a.y_vec.ptr(), yv.data(), c * 2, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
        return;
    }
    mark(ctx, ST_PULLING);
    /* particle-cell tree levels lmin .. top (= lmax - 1), one contiguous u8 buffer */
    const int top = lmax - 1;
    const std::vector<int64_t> off = pct_layout(a);
    Buf pct;
    uint8_t* P = seeded;
    auto D = [&](int l) { return Dims{a.y_num[l], a.x_num[l], a.z_num[l]}; };
    if (!P) {
        pct.reset(ctx->pool, (size_t)off[top + 1]);
        P = pct.as<uint8_t>();
        seed_levels(ctx, a, level_top, a.z_num, P, off);
    }
    /* pulling scheme, top to bottom */
    for (int l = top; l >= lmin; --l) {
        const bool has_parent = (l - 1) >= lmin;
        const bool is_top = (l == top);
        uint8_t* Pc = is_top ? nullptr : P + off[l + 1];
        uint8_t* Pp = has_parent ? P + off[l - 1] : nullptr;
        auto pargs = [&](int zc) {
            const int64_t G = stride_tiles(cdiv(a.y_num[l], 8) * a.x_num[l] * cdiv(a.z_num[l], zc), 256 * 2);
            return PullLevelArgs{P + off[l], D(l), Pc, is_top ? Dims{1, 1, 1} : D(l + 1), Pp, has_parent ? D(l - 1) : Dims{1, 1, 1},
                                 is_top ? 1 : 0, has_parent ? 1 : 0, G};
        };
        const PullLevelArgs p4 = pargs(PULL_ZC);
        if (ctx->recorder) {
            const PullLevelArgs p8 = pargs(8);
            launch_tiles_alts(ctx, {tile_alt<PullLevel<PULL_ZC>>(p4, p4.G, "4 planes per lane"), tile_alt<PullLevel<8>>(p8, p8.G, "8 planes per lane")},
                              0);
        } else {
            launch_tiles<PullLevel<>>(ctx, p4, p4.G);
        }
    }
    const int min_type = par.neighborhood_optimization ? 1 : 2;
    mark(ctx, ST_STRUCTURE);
    /* equivalence optimisation, levels upward in order */
    for (int l = lmin + 1; l <= top; ++l) {
        const int64_t G = stride_tiles(off[l + 1] - off[l], 256 * 4);
        const int i32 = fits32(off[l + 1] - off[l]) ? 1 : 0;
        launch_tiles<Upseed>(ctx, {P + off[l], D(l), P + off[l - 1], D(l - 1), min_type, G, i32 ? FastDiv((uint64_t)a.y_num[l]) : FastDiv(),
                                   i32 ? FastDiv((uint64_t)a.x_num[l]) : FastDiv(), i32},
                             G);
    }
    /* segment table over APR levels lmin..lmax */
    RowTable t;
    std::memset(&t, 0, sizeof(t));
    t.l_min = lmin;
    t.l_max = lmax;
    t.n_levels = lmax - lmin + 1;
    if (t.n_levels > MAXL) throw Error(APRH_ERR_UNSUPPORTED, "too many levels");
    t.min_type = min_type;
    t.y_max = a.y_num[lmax];
    int64_t rp = 0;
    for (int l = lmin; l <= lmax; ++l) {
        const int li = l - lmin;
        const int ml = (l == lmax) ? top : l;   /* pct map providing the cells */
        t.row_begin[li] = (int64_t)a.level_xz_vec[l];
        t.rows_x[li] = a.x_num[l];
        t.map_off[li] = off[ml];
        t.map_y[li] = a.y_num[ml];
        t.map_x[li] = a.x_num[ml];
        t.row_prefix[li] = rp;
        rp += a.x_num[l] * a.z_num[l];
    }
    t.row_prefix[t.n_levels] = rp;
    t.top_z = a.z_num[lmax];
    t.top_pz = a.z_num[top];
    t.n_work = t.row_prefix[t.n_levels - 1] + a.x_num[top] * a.z_num[top];
    /* per-row counts, exclusive scan over rows in particle order, xz_end_vec, y_vec */
    Buf cnt(ctx->pool, (size_t)rp * 4);
    Buf rowoff(ctx->pool, (size_t)rp * 8);
    const int64_t RG = stride_tiles(t.n_work, 16, 16384);   /* 16 lanes per work item, 16 items per block */
    launch_tiles<RowCount>(ctx, {P, t, rp, cnt.as<uint32_t>(), RG}, RG);
    if (ctx->async_mode) {
        /* y_vec and xz_end_vec are preallocated (y_vec for the upper bound P <= N); the count
           stays on the device */
        exclusive_scan_u32_async(ctx, cnt.as<uint32_t>(), rowoff.as<uint64_t>(), rp, ctx->slot_total);
    } else {
        a.total = exclusive_scan_u32(ctx, cnt.as<uint32_t>(), rowoff.as<uint64_t>(), rp);
        a.xz_end.reset(ctx->pool, a.n_xz * 8);
        a.y_vec.reset(ctx->pool, std::max<uint64_t>(a.total, 1) * 2);
    }
    launch_zero(ctx, a.xz_end.ptr(), a.n_xz * 8);
    launch_tiles<RowsXz>(ctx, {rowoff.as<uint64_t>(), cnt.as<uint32_t>(), t, a.xz_end.as<uint64_t>(), rp}, cdiv(rp, 256));
    RowImages ri;
    std::memset(&ri, 0, sizeof(ri));
    SampleReq* sr = ctx->sample_req;
    if (sr && sr->img && !ctx->async_mode && !ctx->recorder && a.total <= sr->capacity) {
        if (sr->top_built) {
            /* level lmax-1 came with the y pass */
            pyramid_levels<uint16_t>(ctx, a, sr->img, sr->pyr.as<uint16_t>(), sr->pyr_off, lmax - 2);
        } else {
            sr->pyr_off = build_pyramid<uint16_t>(ctx, a, sr->img, sr->pyr);
        }
        for (int l = lmin; l <= lmax; ++l) {
            ri.img[l - lmin] = (l == lmax) ? sr->img : sr->pyr.as<uint16_t>() + sr->pyr_off[l];
            ri.y[l - lmin] = a.y_num[l];
        }
        ri.parts = sr->parts;
        sr->done = true;
    }
    /* a wave per work item when the longest rows fill one (1024 cells), else 16 lanes */
    {
        const int64_t FG16 = stride_tiles(t.n_work, FILL_WAVES * 4, 16384), FG64 = stride_tiles(t.n_work, FILL_WAVES, 16384);
        const typename RowFill<16>::Args f16{P, t, rp, rowoff.as<uint64_t>(), a.y_vec.as<uint16_t>(), FG16, ri};
        const typename RowFill<64>::Args f64{P, t, rp, rowoff.as<uint64_t>(), a.y_vec.as<uint16_t>(), FG64, ri};
        const int prefer = t.map_y[t.n_levels - 1] >= 512 ? 1 : 0;
        if (ctx->recorder)
            launch_tiles_alts(ctx, {tile_alt<RowFill<16>>(f16, FG16, "16 lanes per item"), tile_alt<RowFill<64>>(f64, FG64, "64 lanes per item")},
                              prefer);
        else if (prefer)
            launch_tiles<RowFill<64>>(ctx, f64, FG64);
        else
            launch_tiles<RowFill<16>>(ctx, f16, FG16);
    }
    if (ctx->keep_intermediates) {
        a.pct.clear();
        a.pct_dims.assign(3 * (top + 1), 0);
        a.pct.resize(top + 1);
        for (int l = lmin; l <= top; ++l) {
            const size_t nb2 = (size_t)(a.y_num[l] * a.x_num[l] * a.z_num[l]);
            a.pct[l].reset(ctx->pool, nb2);
            HIP_CHECK(hipMemcpyAsync(a.pct[l].ptr(), P + off[l], nb2, hipMemcpyDeviceToDevice, st));
            a.pct_dims[3 * l] = a.y_num[l];
            a.pct_dims[3 * l + 1] = a.x_num[l];
            a.pct_dims[3 * l + 2] = a.z_num[l];
        }
    }
    if (!ctx->async_mode) HIP_CHECK(hipStreamSynchronize(st));
}

/* layout of the mean pyramid of an image over levels lmin .. lmax-1 (PixelData.hpp:1036-1058):
   one buffer, 16-byte aligned levels; per-level element offsets and the total */
std::vector<int64_t> pyramid_layout(const aprh_apr& a, int64_t* total) {
    std::vector<int64_t> off(a.l_max + 1, 0);
    int64_t tot = 0;
    for (int l = a.l_min; l < a.l_max; ++l) {
        off[l] = tot;
        tot += (a.y_num[l] * a.x_num[l] * a.z_num[l] + 7) / 8 * 8;
    }
    *total = tot;
    return off;
}

/* levels `first` .. lmin of the pyramid, each from the level above (the image above lmax-1) */
template <typename T>
void pyramid_levels(aprh_context* ctx, const aprh_apr& a, const T* img, T* Pyr, const std::vector<int64_t>& off, int first) {
    hipStream_t st = ctx->stream;
    const int lmax = a.l_max, lmin = a.l_min;
    for (int l = first; l >= lmin; --l) {
        const T* fine = (l == lmax - 1) ? img : Pyr + off[l + 1];
        const Dims fd{a.y_num[l + 1], a.x_num[l + 1], a.z_num[l + 1]}, cd{a.y_num[l], a.x_num[l], a.z_num[l]};
        T* coarse = Pyr + off[l];
        if (std::is_same<T, uint16_t>::value && fd.y % 8 == 0 && cd.y * 2 == fd.y && ((uintptr_t)fine & 15u) == 0 &&
            ((uintptr_t)coarse & 7u) == 0) {
            const int64_t work = cd.y / 4 * cd.x * cd.z;
            hipLaunchKernelGGL(k_pyr_down_u16v, dim3((unsigned)std::min<int64_t>(cdiv(work, 256), 16384)), dim3(256), 0, st,
                               (const uint16_t*)fine, fd, (uint16_t*)coarse, cd);
        } else {
            hipLaunchKernelGGL(k_pyr_down<T>, dim3(4096), dim3(256), 0, st, fine, fd, coarse, cd);
        }
        HIP_LAUNCH_CHECK();
    }
}

template <typename T>
std::vector<int64_t> build_pyramid(aprh_context* ctx, const aprh_apr& a, const T* img, Buf& pyr) {
    int64_t tot = 0;
    std::vector<int64_t> off = pyramid_layout(a, &tot);
    pyr.reset(ctx->pool, (size_t)std::max<int64_t>(tot, 1) * sizeof(T));
    pyramid_levels<T>(ctx, a, img, pyr.as<T>(), off, a.l_max - 1);
    return off;
}

/* pyr_ext: caller-provided pyramid memory of at least sample_pyramid_bytes() (nothing is
   allocated then); nullptr: the pyramid comes from the context's pool */
template <typename T>
void launch_sample(aprh_context* ctx, const aprh_apr& a, const T* img, T* parts, T* pyr_ext) {
    hipStream_t st = ctx->stream;
    const int lmax = a.l_max, lmin = a.l_min;
    Buf pyr;
    std::vector<int64_t> off;
    T* Pyr = pyr_ext;
    if (pyr_ext) {
        int64_t tot = 0;
        off = pyramid_layout(a, &tot);
        pyramid_levels<T>(ctx, a, img, pyr_ext, off, a.l_max - 1);
    } else {
        off = build_pyramid<T>(ctx, a, img, pyr);
        Pyr = pyr.as<T>();
    }
    LevelPtrs<T> L;
    std::memset(&L, 0, sizeof(L));
    L.n_levels = lmax - lmin + 1;
    int64_t rp = 0;
    for (int l = lmin; l <= lmax; ++l) {
        const int li = l - lmin;
        L.img[li] = (l == lmax) ? img : Pyr + off[l];
        L.y[li] = a.y_num[l];
        L.x[li] = a.x_num[l];
        L.row_begin[li] = (int64_t)a.level_xz_vec[l];
        L.nrows[li] = a.x_num[l] * a.z_num[l];
        L.rows_prefix[li] = rp;
        rp += L.nrows[li];
    }
    L.rows_prefix[L.n_levels] = rp;
    if (a.total > 0) {
        hipLaunchKernelGGL(k_gather<T>, dim3(8192), dim3(256), 0, st, L, a.xz_end.as<uint64_t>(), a.y_vec.as<uint16_t>(), parts, rp);
        HIP_LAUNCH_CHECK();
    }
    HIP_CHECK(hipStreamSynchronize(st));
}

void launch_sample_u16(aprh_context* ctx, const aprh_apr& a, const uint16_t* img, uint16_t* parts, uint16_t* pyr_ext) {
    launch_sample<uint16_t>(ctx, a, img, parts, pyr_ext);
}
void launch_sample_f32(aprh_context* ctx, const aprh_apr& a, const float* img, float* parts) { launch_sample<float>(ctx, a, img, parts, nullptr); }

size_t sample_pyramid_bytes(const aprh_apr& a, size_t elem) {
    int64_t tot = 0;
    pyramid_layout(a, &tot);
    return (size_t)tot * elem;
}

}  // namespace aprh
