/*
 * k_particles.hip -- operations on the particles of a device-resident APR:
 * tree (interior cells), fill_tree_mean, piecewise-constant reconstruction,
 * APR convolution and Richardson-Lucy.
 *
 *  - tree (APR::initialize_tree_dense, APR.hpp:633-689): per-level dense u8
 *    maps; one wave per particle row marks the parents of its particles, the
 *    maps are OR-ed upward through the (0,0,0) child, and the linear tree
 *    access is built by row counts + scan + fill.
 *  - fill_tree_mean (APRTreeNumerics.hpp:63-202): tree levels from the top
 *    down; the children level is scattered into a dense value/type scratch and
 *    each tree node accumulates scale*value/8 over its particle children first
 *    and its tree children second, both in (z, x, y) order -- the order in which
 *    the reference's two passes reach every node -- so the sums are bit-equal.
 *  - level images: the piecewise-constant image at level l is built coarse to
 *    fine (upsample level l-1, then write the particles of level l and, where
 *    finer particles live, the tree means of level l).  It is the isotropic
 *    neighbourhood that APRFilter::convolve_pencil (APRFilter.hpp:1094-1216)
 *    assembles in its ring buffer, and at level l_max + level_delta it is
 *    APRReconstruction::reconstruct_constant (APRReconstruction.hpp:80-120,
 *    :297-394).
 *  - convolution: every particle sums stencil x level image over its reflected
 *    or zero-padded window in the reference's (z, x, y) tap order.
 *  - Richardson-Lucy (APRNumerics.hpp:549-593) with an optional relative-update
 *    stopping rule; the update norms are reduced in a fixed order.
 */
#include <algorithm>
#include <cmath>

#include "aprh_internal.hpp"
#include "tiles.hpp"

namespace aprh {

namespace {

constexpr int MAXL = 24;

/* rows of a linear access (APR or tree) over levels lo..hi */
struct Rows {
    int n_levels, l_lo;
    int64_t prefix[MAXL + 1];     /* first global row id of each level */
    int64_t row_begin[MAXL];      /* index in xz_end of row 0          */
    int64_t x_num[MAXL], y_num[MAXL], z_num[MAXL];
};

Rows make_rows(const std::vector<int64_t>& yn, const std::vector<int64_t>& xn, const std::vector<int64_t>& zn,
               const std::vector<uint64_t>& level_xz, int lo, int hi) {
    Rows r;
    std::memset(&r, 0, sizeof(r));
    r.n_levels = std::max(hi - lo + 1, 0);
    r.l_lo = lo;
    if (r.n_levels > MAXL) throw Error(APRH_ERR_UNSUPPORTED, "too many levels");
    int64_t p = 0;
    for (int l = lo; l <= hi; ++l) {
        const int i = l - lo;
        r.prefix[i] = p;
        r.row_begin[i] = (int64_t)level_xz[l];
        r.x_num[i] = xn[l];
        r.y_num[i] = yn[l];
        r.z_num[i] = zn[l];
        p += xn[l] * zn[l];
    }
    r.prefix[r.n_levels] = p;
    return r;
}

__device__ __forceinline__ int row_level(const Rows& R, int64_t r) {
    int li = 0;
    while (li + 1 < R.n_levels && r >= R.prefix[li + 1]) ++li;
    return li;
}

__device__ __forceinline__ uint16_t f2u16(float v) {
    /* g++/x86-64 float->uint16 conversion: truncation through int32 */
    if (!(v > -2147483648.0f && v < 2147483648.0f)) return 0;
    return (uint16_t)(int32_t)v;
}

template <typename D> struct Cvt;
template <> struct Cvt<float> {
    static __device__ __forceinline__ float from(float v) { return v; }
    static __device__ __forceinline__ float from(uint16_t v) { return (float)v; }
};
template <> struct Cvt<uint16_t> {
    static __device__ __forceinline__ uint16_t from(float v) { return f2u16(v); }
    static __device__ __forceinline__ uint16_t from(uint16_t v) { return v; }
};

/* ---------------- tree build ---------------- */
/* APR particle (level l) marks tree cell (l-1, y/2, x/2, z/2) */
__global__ __launch_bounds__(256) void k_mark_parents(Rows R, const uint64_t* __restrict__ xz_end, const uint16_t* __restrict__ y_vec,
                                                      uint8_t* const* maps, int64_t total_rows) {
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / 64;
    for (int64_t r = wid; r < total_rows; r += nw) {
        const int li = row_level(R, r);
        const int level = R.l_lo + li;
        const int64_t row = r - R.prefix[li];
        const int64_t z = row / R.x_num[li], x = row % R.x_num[li];
        const uint64_t b = xz_end[R.row_begin[li] + row - 1], e = xz_end[R.row_begin[li] + row];
        uint8_t* m = maps[level - 1];
        const int64_t pyn = (R.y_num[li] + 1) / 2, pxn = (R.x_num[li] + 1) / 2;
        uint8_t* prow = m + ((z / 2) * pxn + (x / 2)) * pyn;
        for (uint64_t i = b + lane; i < e; i += 64) prow[y_vec[i] / 2] = 1;
    }
}

/* tree[l](y,x,z) |= tree[l+1](2y,2x,2z) */
__global__ void k_or_up(uint8_t* m, int64_t yn, int64_t xn, int64_t zn, const uint8_t* __restrict__ c, int64_t cyn, int64_t cxn) {
    const size_t n = (size_t)(yn * xn * zn);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int64_t y = (int64_t)(i % (size_t)yn);
        const int64_t r = (int64_t)(i / (size_t)yn);
        const int64_t x = r % xn, z = r / xn;
        m[i] = m[i] | c[((2 * z) * cxn + 2 * x) * cyn + 2 * y];
    }
}

/* per-row count of a 0/1 map row (LinearAccess::initialize_tree_access_dense, LinearAccess.hpp:75-90):
   one wave per row */
__global__ __launch_bounds__(256) void k_map_row_count(Rows R, const uint8_t* const* maps, uint32_t* cnt, int64_t total_rows) {
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / 64;
    for (int64_t r = wid; r < total_rows; r += nw) {
        const int li = row_level(R, r);
        const int64_t row = r - R.prefix[li];
        const int64_t yn = R.y_num[li];
        const uint8_t* p = maps[R.l_lo + li] + row * yn;
        uint32_t c = 0;
        for (int64_t y = lane; y < yn; y += 64) c += p[y];
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if (lane == 0) cnt[r] = c;
    }
}

__global__ __launch_bounds__(256) void k_map_row_fill(Rows R, const uint8_t* const* maps, const uint64_t* __restrict__ xz_end,
                                                      uint16_t* y_vec, int64_t total_rows) {
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / 64;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int64_t r = wid; r < total_rows; r += nw) {
        const int li = row_level(R, r);
        const int64_t row = r - R.prefix[li];
        const int64_t yn = R.y_num[li];
        const uint8_t* p = maps[R.l_lo + li] + row * yn;
        uint64_t o = xz_end[R.row_begin[li] + row - 1];
        for (int64_t y0 = 0; y0 < yn; y0 += 64) {
            const int64_t y = y0 + lane;
            const bool on = y < yn && p[y];
            const uint64_t b = __ballot(on);
            if (on) y_vec[o + __popcll(b & lt)] = (uint16_t)y;
            o += __popcll(b);
        }
    }
}

/* xz_end[row_begin + row] = inclusive scan value of global row r */
__global__ void k_rows_to_xz(Rows R, const uint64_t* __restrict__ excl, const uint32_t* __restrict__ cnt, uint64_t* xz_end,
                             int64_t total_rows) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= total_rows) return;
    const int li = row_level(R, r);
    xz_end[R.row_begin[li] + (r - R.prefix[li])] = excl[r] + cnt[r];
}

/* ---------------- dense level images ---------------- */
/* values (and optionally a type tag) of every particle / tree node of one level into a dense
   image: wave `wid` of `nw` takes rows wid, wid + nw, ..., its lanes the row's entries */
template <typename T, typename D>
__device__ __forceinline__ void scatter_rows_body(const Rows& R, int li, const uint64_t* __restrict__ xz_end,
                                                  const uint16_t* __restrict__ y_vec, const T* __restrict__ vals, D* V, uint8_t* Ty,
                                                  uint8_t type, int64_t wid, int64_t nw, int lane) {
    const int64_t nrows = R.x_num[li] * R.z_num[li];
    for (int64_t row = wid; row < nrows; row += nw) {
        const uint64_t b = xz_end[R.row_begin[li] + row - 1], e = xz_end[R.row_begin[li] + row];
        D* vr = V + row * R.y_num[li];
        uint8_t* tr = Ty ? Ty + row * R.y_num[li] : nullptr;
        for (uint64_t i = b + lane; i < e; i += 64) {
            const uint16_t y = y_vec[i];
            vr[y] = Cvt<D>::from(vals[i]);
            if (tr) tr[y] = type;
        }
    }
}
template <typename T, typename D>
__global__ __launch_bounds__(256) void k_scatter_rows(Rows R, int li, const uint64_t* __restrict__ xz_end,
                                                      const uint16_t* __restrict__ y_vec, const T* __restrict__ vals, D* V,
                                                      uint8_t* Ty, uint8_t type) {
    scatter_rows_body<T, D>(R, li, xz_end, y_vec, vals, V, Ty, type, ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64,
                            ((int64_t)gridDim.x * blockDim.x) / 64, threadIdx.x & 63);
}

/* one tree level of fill_tree_mean: children (level t+1) are in the dense V/Ty scratch */
__device__ __forceinline__ void tree_mean_level_body(const Rows& TR, int ti, const uint64_t* __restrict__ txz,
                                                     const uint16_t* __restrict__ ty, const float* __restrict__ V,
                                                     const uint8_t* __restrict__ Ty, int64_t cyn, int64_t cxn, int64_t czn, float* tree,
                                                     int64_t wid, int64_t nw, int lane) {
    const int64_t pxn = TR.x_num[ti], pzn = TR.z_num[ti], pyn = TR.y_num[ti];
    const int64_t nrows = pxn * pzn;
    for (int64_t row = wid; row < nrows; row += nw) {
        const int64_t zp = row / pxn, xp = row % pxn;
        const uint64_t b = txz[TR.row_begin[ti] + row - 1], e = txz[TR.row_begin[ti] + row];
        /* boundary scale factors (APRTreeNumerics.hpp:99-113) */
        float sxz = (float)((((2 * pxn != cxn) && (xp == pxn - 1)) + ((2 * pzn != czn) && (zp == pzn - 1))) * 2);
        if (sxz == 0) sxz = 1;
        const float syxz = (2 * pyn != cyn) ? sxz * 2 : sxz;
        const int nz = (int)min((int64_t)2, czn - 2 * zp), nx = (int)min((int64_t)2, cxn - 2 * xp);
        for (uint64_t i = b + lane; i < e; i += 64) {
            const int64_t yp = ty[i];
            const float s = (yp == pyn - 1) ? syxz : sxz;
            const int ny = (int)min((int64_t)2, cyn - 2 * yp);
            float acc = 0.0f;
            for (int pass = 1; pass <= 2; ++pass) {
                for (int a = 0; a < nz; ++a)
                    for (int bb = 0; bb < nx; ++bb)
                        for (int c = 0; c < ny; ++c) {
                            const int64_t ci = ((2 * zp + a) * cxn + (2 * xp + bb)) * cyn + (2 * yp + c);
                            if (Ty[ci] == pass) acc = s * V[ci] / 8.0f + acc;
                        }
            }
            tree[i] = acc;
        }
    }
}
__global__ __launch_bounds__(256) void k_tree_mean_level(Rows TR, int ti, const uint64_t* __restrict__ txz, const uint16_t* __restrict__ ty,
                                                         const float* __restrict__ V, const uint8_t* __restrict__ Ty, int64_t cyn,
                                                         int64_t cxn, int64_t czn, float* tree) {
    tree_mean_level_body(TR, ti, txz, ty, V, Ty, cyn, cxn, czn, tree, ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64,
                         ((int64_t)gridDim.x * blockDim.x) / 64, threadIdx.x & 63);
}

/* the block's writes to global memory visible to its own later reads: a workgroup-scope fence
   and barrier (__syncthreads).  All waves of a workgroup run on one CU and share its vector L1,
   so no cache maintenance is needed at this scope (an agent-scope fence would write back L2) */
__device__ __forceinline__ void block_sync_global() { __syncthreads(); }

/* the small tree levels t_hi .. 0 in one block of 1024 threads (each level's rows are few, so
   separate launches would cost more than their work): per level, clear the type tags of the
   children level, scatter its particles and tree nodes, take the means; the same bodies as the
   per-level kernels, so the results are identical */
template <typename T>
__global__ __launch_bounds__(1024) void k_tree_small(Rows AR, Rows TR, int t_hi, int l_min, int t_l_max, const uint64_t* __restrict__ xz_end,
                                                     const uint16_t* __restrict__ y_vec, const T* __restrict__ parts,
                                                     const uint64_t* __restrict__ txz, const uint16_t* __restrict__ ty, float* tree, float* V,
                                                     uint8_t* Ty) {
    const int64_t wid = threadIdx.x / 64, nw = blockDim.x / 64;
    const int lane = threadIdx.x & 63;
    for (int t = t_hi; t >= 0; --t) {
        const int c = t + 1;
        const int64_t cyn = TR.y_num[c], cxn = TR.x_num[c], czn = TR.z_num[c];
        const int64_t n = cyn * cxn * czn;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) Ty[i] = 0;
        block_sync_global();
        if (c >= l_min) scatter_rows_body<T, float>(AR, c - l_min, xz_end, y_vec, parts, V, Ty, (uint8_t)1, wid, nw, lane);
        if (c <= t_l_max) scatter_rows_body<float, float>(TR, c, txz, ty, (const float*)tree, V, Ty, (uint8_t)2, wid, nw, lane);
        block_sync_global();
        tree_mean_level_body(TR, t, txz, ty, V, Ty, cyn, cxn, czn, tree, wid, nw, lane);
        block_sync_global();
    }
}

/* L(z, x, y) = C(z/2, x/2, y/2) */
template <typename D>
__global__ void k_upsample(const D* __restrict__ C, int64_t cyn, int64_t cxn, D* L, int64_t yn, int64_t xn, int64_t zn) {
    const size_t n = (size_t)(yn * xn * zn);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int64_t y = (int64_t)(i % (size_t)yn);
        const int64_t r = (int64_t)(i / (size_t)yn);
        const int64_t x = r % xn, z = r / xn;
        L[i] = C[((z / 2) * cxn + (x / 2)) * cyn + (y / 2)];
    }
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

/* a stored structure is well formed: rows non-decreasing within [0, total], y strictly
   increasing inside a row and below the level's y_num */
__global__ void k_validate_rows(Rows R, const uint64_t* __restrict__ xz_end, const uint16_t* __restrict__ y_vec, int64_t total_rows,
                                uint64_t total, int* err) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= total_rows) return;
    const int li = row_level(R, r);
    const int64_t row = r - R.prefix[li];
    const uint64_t b = xz_end[R.row_begin[li] + row - 1], e = xz_end[R.row_begin[li] + row];
    if (e < b || e > total) {
        atomicAdd(err, 1);
        return;
    }
    int bad = 0;
    for (uint64_t i = b; i < e; ++i) {
        if ((int64_t)y_vec[i] >= R.y_num[li]) bad = 1;
        if (i > b && y_vec[i] <= y_vec[i - 1]) bad = 1;
    }
    if (bad) atomicAdd(err, 1);
}

const dim3 kGrid(4096), kBlock(256);

size_t level_elems(const aprh_apr& a, int l) { return (size_t)(a.y_num[l] * a.x_num[l] * a.z_num[l]); }

/* build the level image of level `lv` into `cur` from the level image of lv-1 in `prev` (or nothing) */
template <typename T, typename D>
void build_level(aprh_context* ctx, const aprh_apr& a, const Rows& AR, const Rows& TR, int lv, const D* prev, D* cur, const T* parts,
                 const float* tree, bool with_tree) {
    hipStream_t st = ctx->stream;
    if (prev) {
        hipLaunchKernelGGL(k_upsample<D>, kGrid, kBlock, 0, st, prev, a.y_num[lv - 1], a.x_num[lv - 1], cur, a.y_num[lv], a.x_num[lv],
                           a.z_num[lv]);
        HIP_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL((k_scatter_rows<T, D>), kGrid, kBlock, 0, st, AR, lv - a.l_min, a.xz_end.as<uint64_t>(), a.y_vec.as<uint16_t>(), parts,
                       cur, (uint8_t*)nullptr, (uint8_t)0);
    HIP_LAUNCH_CHECK();
    if (with_tree && lv <= a.t_l_max && a.t_total > 0) {
        hipLaunchKernelGGL((k_scatter_rows<float, D>), kGrid, kBlock, 0, st, TR, lv, a.t_xz_end.as<uint64_t>(), a.t_y_vec.as<uint16_t>(),
                           tree, cur, (uint8_t*)nullptr, (uint8_t)0);
        HIP_LAUNCH_CHECK();
    }
}

}  // namespace

/* ======================================================================== */

void validate_structure(aprh_context* ctx, const aprh_apr& a) {
    hipStream_t st = ctx->stream;
    Rows R = make_rows(a.y_num, a.x_num, a.z_num, a.level_xz_vec, a.l_min, a.l_max);
    const int64_t nrows = R.prefix[R.n_levels];
    Buf err(ctx->pool, 4);
    HIP_CHECK(hipMemsetAsync(err.ptr(), 0, 4, st));
    if (nrows > 0) {
        hipLaunchKernelGGL(k_validate_rows, dim3((unsigned)cdiv(nrows, 256)), kBlock, 0, st, R, a.xz_end.as<uint64_t>(),
                           a.y_vec.as<uint16_t>(), nrows, a.total, err.as<int>());
        HIP_LAUNCH_CHECK();
    }
    int h_err = 0;
    uint64_t first = 0, last = 0;
    HIP_CHECK(hipMemcpyAsync(&h_err, err.ptr(), 4, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(&first, a.xz_end.ptr(), 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(&last, a.xz_end.as<uint64_t>() + a.n_xz - 1, 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (h_err || first != 0 || last != a.total)
        throw Error(APRH_ERR_INVALID_ARGUMENT, "malformed APR structure (xz_end_vec / y_vec do not describe the given dimensions)");
}

void ensure_tree(aprh_context* ctx, aprh_apr& a) {
    if (a.tree_built) return;
    hipStream_t st = ctx->stream;
    /* tree geometry (GenInfo::init_tree, GenInfo.hpp:68-99) */
    const int64_t y = a.org[0], x = a.org[1], z = a.org[2];
    const int64_t md = std::max(std::max(y, x), z);
    const int lm = (int)std::ceil(std::log2((double)md));
    a.t_l_max = std::max(lm - 1, 0);
    if (a.t_l_max + 1 > a.l_max) throw Error(APRH_ERR_UNSUPPORTED, "tree of a degenerate APR");
    a.t_y.assign(a.t_l_max + 1, 0);
    a.t_x.assign(a.t_l_max + 1, 0);
    a.t_z.assign(a.t_l_max + 1, 0);
    for (int l = 0; l <= a.t_l_max; ++l) {
        const double cs = std::round(std::pow(2.0, (double)(a.t_l_max - l + 1)));
        a.t_y[l] = (int64_t)std::ceil((double)y / cs);
        a.t_x[l] = (int64_t)std::ceil((double)x / cs);
        a.t_z[l] = (int64_t)std::ceil((double)z / cs);
    }
    a.t_level_xz.assign(a.t_l_max + 2, 0);
    uint64_t ct = 1;
    a.t_level_xz[0] = 1;
    for (int i = 0; i <= a.t_l_max; ++i) {
        ct += (uint64_t)(a.t_x[i] * a.t_z[i]);
        a.t_level_xz[i + 1] = ct;
    }
    a.t_n_xz = ct;
    /* dense maps, one byte per tree cell */
    std::vector<int64_t> off(a.t_l_max + 2, 0);
    for (int l = 0; l <= a.t_l_max; ++l) off[l + 1] = off[l] + a.t_y[l] * a.t_x[l] * a.t_z[l];
    Buf maps(ctx->pool, (size_t)std::max<int64_t>(off[a.t_l_max + 1], 1));
    HIP_CHECK(hipMemsetAsync(maps.ptr(), 0, (size_t)off[a.t_l_max + 1], st));
    std::vector<uint8_t*> mp(MAXL, nullptr);
    for (int l = 0; l <= a.t_l_max; ++l) mp[l] = maps.as<uint8_t>() + off[l];
    Buf d_mp(ctx->pool, MAXL * sizeof(uint8_t*));
    HIP_CHECK(hipMemcpyAsync(d_mp.ptr(), mp.data(), MAXL * sizeof(uint8_t*), hipMemcpyHostToDevice, st));
    if (a.total > 0) {
        Rows R = make_rows(a.y_num, a.x_num, a.z_num, a.level_xz_vec, a.l_min, a.l_max);
        hipLaunchKernelGGL(k_mark_parents, dim3(8192), kBlock, 0, st, R, a.xz_end.as<uint64_t>(), a.y_vec.as<uint16_t>(),
                           (uint8_t* const*)d_mp.ptr(), R.prefix[R.n_levels]);
        HIP_LAUNCH_CHECK();
    }
    for (int l = a.t_l_max - 1; l >= 0; --l) {
        hipLaunchKernelGGL(k_or_up, kGrid, kBlock, 0, st, mp[l], a.t_y[l], a.t_x[l], a.t_z[l], mp[l + 1], a.t_y[l + 1], a.t_x[l + 1]);
        HIP_LAUNCH_CHECK();
    }
    /* linear access of the tree */
    Rows TR = make_rows(a.t_y, a.t_x, a.t_z, a.t_level_xz, 0, a.t_l_max);
    const int64_t nrows = TR.prefix[TR.n_levels];
    Buf cnt(ctx->pool, (size_t)std::max<int64_t>(nrows, 1) * 4);
    Buf excl(ctx->pool, (size_t)std::max<int64_t>(nrows, 1) * 8);
    a.t_xz_end.reset(ctx->pool, a.t_n_xz * 8);
    HIP_CHECK(hipMemsetAsync(a.t_xz_end.ptr(), 0, a.t_n_xz * 8, st));
    uint64_t total = 0;
    if (nrows > 0) {
        hipLaunchKernelGGL(k_map_row_count, dim3((unsigned)std::min<int64_t>(cdiv(nrows, 4), 16384)), kBlock, 0, st, TR,
                           (const uint8_t* const*)d_mp.ptr(), cnt.as<uint32_t>(), nrows);
        HIP_LAUNCH_CHECK();
        total = exclusive_scan_u32(ctx, cnt.as<uint32_t>(), excl.as<uint64_t>(), nrows);
        hipLaunchKernelGGL(k_rows_to_xz, dim3((unsigned)cdiv(nrows, 256)), kBlock, 0, st, TR, excl.as<uint64_t>(), cnt.as<uint32_t>(),
                           a.t_xz_end.as<uint64_t>(), nrows);
        HIP_LAUNCH_CHECK();
    }
    a.t_total = total;
    a.t_y_vec.reset(ctx->pool, std::max<uint64_t>(total, 1) * 2);
    if (nrows > 0 && total > 0) {
        hipLaunchKernelGGL(k_map_row_fill, dim3((unsigned)std::min<int64_t>(cdiv(nrows, 4), 16384)), kBlock, 0, st, TR,
                           (const uint8_t* const*)d_mp.ptr(), a.t_xz_end.as<uint64_t>(), a.t_y_vec.as<uint16_t>(), nrows);
        HIP_LAUNCH_CHECK();
    }
    HIP_CHECK(hipStreamSynchronize(st));
    a.tree_built = true;
}

template <typename T>
void tree_mean(aprh_context* ctx, aprh_apr& a, const T* parts, float* tree) {
    ensure_tree(ctx, a);
    hipStream_t st = ctx->stream;
    if (a.t_total == 0) return;
    Rows AR = make_rows(a.y_num, a.x_num, a.z_num, a.level_xz_vec, a.l_min, a.l_max);
    Rows TR = make_rows(a.t_y, a.t_x, a.t_z, a.t_level_xz, 0, a.t_l_max);
    /* the top tree level straight from the particle rows of l_max when they have the regular
       shape; the levels below through dense scratch of the children level */
    const bool top = tree_mean_top<T>(ctx, a, parts, tree);
    const int t_first = top ? a.t_l_max - 1 : a.t_l_max;
    if (t_first < 0) return;
    const size_t maxn = level_elems(a, t_first + 1);
    Buf V(ctx->pool, maxn * 4), Ty(ctx->pool, maxn);
    /* tree levels whose children level has at most 256 rows go to one single-block launch */
    int t_small = -1;
    for (int t = t_first; t >= 0; --t)
        if (a.x_num[t + 1] * a.z_num[t + 1] <= 256) {
            t_small = t;
            break;
        }
    for (int t = t_first; t > t_small; --t) {
        const int c = t + 1;   /* children level; APR and tree levels share the level geometry */
        const int64_t cyn = a.y_num[c], cxn = a.x_num[c], czn = a.z_num[c];
        HIP_CHECK(hipMemsetAsync(Ty.ptr(), 0, level_elems(a, c), st));
        if (c >= a.l_min) {
            hipLaunchKernelGGL((k_scatter_rows<T, float>), kGrid, kBlock, 0, st, AR, c - a.l_min, a.xz_end.as<uint64_t>(),
                               a.y_vec.as<uint16_t>(), parts, V.as<float>(), Ty.as<uint8_t>(), (uint8_t)1);
            HIP_LAUNCH_CHECK();
        }
        if (c <= a.t_l_max) {
            hipLaunchKernelGGL((k_scatter_rows<float, float>), kGrid, kBlock, 0, st, TR, c, a.t_xz_end.as<uint64_t>(),
                               a.t_y_vec.as<uint16_t>(), (const float*)tree, V.as<float>(), Ty.as<uint8_t>(), (uint8_t)2);
            HIP_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_tree_mean_level, kGrid, kBlock, 0, st, TR, t, a.t_xz_end.as<uint64_t>(), a.t_y_vec.as<uint16_t>(),
                           V.as<float>(), Ty.as<uint8_t>(), cyn, cxn, czn, tree);
        HIP_LAUNCH_CHECK();
    }
    if (t_small >= 0) {
        hipLaunchKernelGGL((k_tree_small<T>), dim3(1), dim3(1024), 0, st, AR, TR, t_small, a.l_min, a.t_l_max, a.xz_end.as<uint64_t>(),
                           a.y_vec.as<uint16_t>(), parts, a.t_xz_end.as<uint64_t>(), a.t_y_vec.as<uint16_t>(), tree, V.as<float>(),
                           Ty.as<uint8_t>());
        HIP_LAUNCH_CHECK();
    }
}

std::vector<HostStencil> downsampled_stencils(const HostStencil& in, int nlevels, bool normalize) {
    /* APRStencil::downsample_stencil (APRStencil.hpp:12-78): per output tap the area-overlap
       weighted sum over input taps, visited in (z, x, y) input order */
    std::vector<HostStencil> out((size_t)std::max(nlevels, 1));
    out[0] = in;
    if (normalize) {
        /* PixelData::normalize (PixelData.hpp:771-781) */
        float sum = 0;
        for (float v : out[0].w) sum += v;
        for (auto& v : out[0].w) v /= sum;
    }
    const int nd = (in.sz > 1) + (in.sx > 1) + (in.sy > 1);
    for (int d = 1; d < nlevels; ++d) {
        const int step = (int)std::pow(2.0f, (float)d);
        const int factor = (int)std::pow((float)step, (float)nd);
        auto ds = [&](int n) { return n > 1 ? std::max(2 * (((n - 1) / 2 + step - 1) / step) + 1, 3) : 1; };
        HostStencil& o = out[(size_t)d];
        o.sy = ds(in.sy);
        o.sx = ds(in.sx);
        o.sz = ds(in.sz);
        o.w.assign((size_t)o.sy * o.sx * o.sz, 0.0f);
        const int oz = (o.sz / 2) * step - in.sz / 2, ox = (o.sx / 2) * step - in.sx / 2, oy = (o.sy / 2) * step - in.sy / 2;
        auto overlap = [&](int dcell, int off_i) {
            return (float)std::max(std::min(dcell * step + step, off_i + step) - std::max(dcell * step, off_i), 0);
        };
        for (int dz = 0; dz < o.sz; ++dz)
            for (int dx = 0; dx < o.sx; ++dx)
                for (int dy = 0; dy < o.sy; ++dy) {
                    float acc = 0.0f;
                    for (int iz = 0; iz < in.sz; ++iz) {
                        const float wz = overlap(dz, oz + iz);
                        for (int ix = 0; ix < in.sx; ++ix) {
                            const float wx = overlap(dx, ox + ix);
                            for (int iy = 0; iy < in.sy; ++iy) {
                                const float wy = overlap(dy, oy + iy);
                                acc += (wx * wy * wz) * in.w[((size_t)iz * in.sx + ix) * in.sy + iy];
                            }
                        }
                    }
                    o.w[((size_t)dz * o.sx + dx) * o.sy + dy] = acc;
                }
        float sum = 0;
        for (auto& v : o.w) {
            v /= (float)factor;
            sum += v;
        }
        if (normalize) {
            const float nf = 1.0f / sum;
            for (auto& v : o.w) v *= nf;
        }
    }
    return out;
}

/* 1-D factors (fy, fx, fz) with fz (x) fx (x) fy == w to within 1e-6 of max|w|, if they exist */
bool rank1_factors(const HostStencil& s, std::vector<float>& fy, std::vector<float>& fx, std::vector<float>& fz) {
    size_t c = 0;
    double mx = 0;
    for (size_t i = 0; i < s.w.size(); ++i)
        if (std::fabs((double)s.w[i]) > mx) {
            mx = std::fabs((double)s.w[i]);
            c = i;
        }
    if (mx == 0) return false;
    const int cy = (int)(c % s.sy), cx = (int)((c / s.sy) % s.sx), cz = (int)(c / ((size_t)s.sy * s.sx));
    auto W = [&](int z, int x, int y) { return (double)s.w[((size_t)z * s.sx + x) * s.sy + y]; };
    const double wc = W(cz, cx, cy);
    std::vector<double> dy(s.sy), dx(s.sx), dz(s.sz);
    for (int y = 0; y < s.sy; ++y) dy[y] = W(cz, cx, y) / wc;
    for (int x = 0; x < s.sx; ++x) dx[x] = W(cz, x, cy) / wc;
    for (int z = 0; z < s.sz; ++z) dz[z] = W(z, cx, cy);
    for (int z = 0; z < s.sz; ++z)
        for (int x = 0; x < s.sx; ++x)
            for (int y = 0; y < s.sy; ++y)
                if (std::fabs(dz[z] * dx[x] * dy[y] - W(z, x, y)) > 1e-6 * mx) return false;
    fy.assign(dy.begin(), dy.end());
    fx.assign(dx.begin(), dx.end());
    fz.assign(dz.begin(), dz.end());
    return true;
}

template <typename T>
void convolve(aprh_context* ctx, aprh_apr& a, const std::vector<HostStencil>& stencils, const T* in, float* out, bool reflect,
              const float* tree_in) {
    if (stencils.empty()) throw Error(APRH_ERR_INVALID_ARGUMENT, "no stencils");
    if (a.total == 0) return;
    ensure_tree(ctx, a);
    Buf tree_buf;
    const float* tree_vals = tree_in;
    if (!tree_vals) {
        /* APRFilter.hpp:1088: tree means of the input */
        tree_buf.reset(ctx->pool, std::max<uint64_t>(a.t_total, 1) * 4);
        tree_mean(ctx, a, in, tree_buf.as<float>());
        tree_vals = tree_buf.as<float>();
    }
    conv_levels<T>(ctx, a, stencils, ctx->conv_mode == APRH_CONV_SEPARABLE, in, tree_vals, out, reflect, 0, nullptr, false, nullptr,
                   nullptr);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

template <typename T, typename U>
void reconstruct(aprh_context* ctx, aprh_apr& a, const T* parts, int level, U* img) {
    if (level < a.l_min || level > a.l_max) throw Error(APRH_ERR_INVALID_ARGUMENT, "reconstruction level out of range");
    hipStream_t st = ctx->stream;
    const bool with_tree = level < a.l_max;
    Buf tree;
    Rows TR;
    std::memset(&TR, 0, sizeof(TR));
    if (with_tree) {
        ensure_tree(ctx, a);
        tree.reset(ctx->pool, std::max<uint64_t>(a.t_total, 1) * 4);
        tree_mean(ctx, a, parts, tree.as<float>());
        TR = make_rows(a.t_y, a.t_x, a.t_z, a.t_level_xz, 0, a.t_l_max);
    }
    Rows AR = make_rows(a.y_num, a.x_num, a.z_num, a.level_xz_vec, a.l_min, a.l_max);
    const size_t maxn = level > a.l_min ? level_elems(a, level - 1) : 1;
    Buf S0(ctx->pool, maxn * sizeof(U)), S1(ctx->pool, maxn * sizeof(U));
    U* prev = nullptr;
    U* bufs[2] = {S0.as<U>(), S1.as<U>()};
    for (int l = a.l_min; l <= level; ++l) {
        U* cur = (l == level) ? img : bufs[l & 1];
        build_level<T, U>(ctx, a, AR, TR, l, prev, cur, parts, tree.as<float>(), with_tree && l == level);
        prev = cur;
    }
    HIP_CHECK(hipStreamSynchronize(st));
}

namespace {
/* value v over n u16 elements (grid-stride over a virtual grid of G blocks) */
struct FillU16 {
    struct Args {
        uint16_t* p;
        uint64_t n;
        uint16_t v;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "fill_u16"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        for (uint64_t i = (uint64_t)t * blockDim.x + threadIdx.x; i < a.n; i += (uint64_t)a.G * blockDim.x) a.p[i] = a.v;
    }
};
}  // namespace

/* APRReconstruction::reconstruct_level (APRReconstruction.hpp:131-171): every pixel takes the
   level of the particle covering it, which is the constant reconstruction of particles whose
   values are their levels (particles are stored level by level, so each level is one range).
   At a coarser level L (the patch form, :397-490) a pixel covered by finer particles takes L
   from the tree: the particles carry min(level, L), so every tree mean there is exactly L. */
void reconstruct_level(aprh_context* ctx, aprh_apr& a, int level, void* img, int idt) {
    Buf lev(ctx->pool, std::max<uint64_t>(a.total, 1) * 2);
    for (int l = a.l_min; l <= a.l_max; ++l) {
        const uint64_t r0 = a.level_xz_vec[l], nrows = (uint64_t)(a.x_num[l] * a.z_num[l]);
        if (nrows == 0) continue;
        uint64_t b = 0, e = 0;
        if (r0 > 0) HIP_CHECK(hipMemcpyAsync(&b, a.xz_end.as<uint64_t>() + r0 - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipMemcpyAsync(&e, a.xz_end.as<uint64_t>() + r0 + nrows - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (e > b) {
            const int64_t G = stride_tiles((int64_t)(e - b), 256 * 4);
            launch_tiles<FillU16>(ctx, {lev.as<uint16_t>() + b, e - b, (uint16_t)std::min(l, level), G}, G);
        }
    }
    reconstruct_any(ctx, a, lev.ptr(), APRH_DTYPE_U16, level, img, idt);
}

namespace {
/* the patch out of a level image: out[z, x, y] = full[(z + zb) >> s, (x + xb) >> s, (y + yb) >> s]
   (s > 0: nearest-neighbour upsampling by 2^s), y fastest in both */
template <typename U>
struct CropPatch {
    struct Args {
        const U* full;
        U* out;
        int64_t fy, fx;
        int64_t py, px, n;
        int64_t yb, xb, zb;
        int s;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "crop_patch"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        for (int64_t i = t * blockDim.x + threadIdx.x; i < a.n; i += a.G * blockDim.x) {
            const int64_t y = i % a.py, r = i / a.py;
            const int64_t x = r % a.px, z = r / a.px;
            a.out[i] = a.full[(((z + a.zb) >> a.s) * a.fx + ((x + a.xb) >> a.s)) * a.fy + ((y + a.yb) >> a.s)];
        }
    }
};
}  // namespace

namespace {
/* |grad| from the per-axis gradients, in APRNumerics::gradient_magnitude_*'s order
   (APRNumerics.hpp:216-283): g_y^2, + g_x^2, + g_z^2 (absent axes skipped), sqrtf */
struct GradMag {
    struct Args {
        float* g;          /* in: g_y, out: magnitude */
        const float* gx;   /* nullptr: no x axis */
        const float* gz;   /* nullptr: no z axis */
        uint64_t n;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "grad_mag"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        for (uint64_t i = (uint64_t)t * blockDim.x + threadIdx.x; i < a.n; i += (uint64_t)a.G * blockDim.x) {
            float m = a.g[i] * a.g[i];
            if (a.gx) m = m + a.gx[i] * a.gx[i];
            if (a.gz) m = m + a.gz[i] * a.gz[i];
            a.g[i] = sqrtf(m);
        }
    }
};
}  // namespace

namespace {
/* APRNumerics::face_neighbour_filter (APRNumerics.hpp:305-355) along one dimension: one lane
   per particle (a wave per row).  The neighbours of each face come from the random iterator's
   search (APRIterator.hpp:283-394, RandomAccess.hpp:92-170): the same-level cell, else (above
   l_min) the parent-level cell containing the face neighbour, else (below l_max) the four
   children on the face, counted where they are particles provided the first one is.  Cells
   are found by binary search in their row; coordinates follow the reference's unsigned 16-bit
   arithmetic, and the four l_max rows under the first parent row read row (0, 0), as the
   reference's random access resolves them. */
struct FaceFilter {
    struct Args {
        Rows R;
        const uint64_t* xz_end;
        const uint16_t* y_vec;
        const void* in;
        int in_u16;
        float* out;
        float f0, f1, f2;
        int dim;
        int64_t total_rows;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "face_neighbour_filter"; }

    static __device__ __forceinline__ float val(const Args& a, int64_t i) {
        return a.in_u16 ? (float)((const uint16_t*)a.in)[i] : ((const float*)a.in)[i];
    }
    /* particle index at (level, z, x, y) or -1 */
    static __device__ __forceinline__ int64_t find(const Args& a, int level, int64_t z, int64_t x, int64_t y) {
        const Rows& R = a.R;
        const int li = level - R.l_lo;
        if (li < 0 || li >= R.n_levels) return -1;
        if (z < 0 || x < 0 || y < 0 || z >= R.z_num[li] || x >= R.x_num[li] || y >= R.y_num[li]) return -1;
        if (li == R.n_levels - 1 && z < 2 && x < 2) {
            z = 0;
            x = 0;
        }
        const int64_t rid = R.row_begin[li] + z * R.x_num[li] + x;
        uint64_t lo = a.xz_end[rid - 1], hi = a.xz_end[rid];
        const uint64_t e = hi;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if ((int64_t)a.y_vec[mid] < y) lo = mid + 1;
            else hi = mid;
        }
        return (lo < e && (int64_t)a.y_vec[lo] == y) ? (int64_t)lo : -1;
    }
    static __device__ __forceinline__ int64_t u16c(int64_t v) { return (int64_t)(uint16_t)v; }

    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        const int lane = threadIdx.x & 63;
        const int64_t wpb = blockDim.x / 64;
        const Rows& R = a.R;
        const int l_min = R.l_lo, l_max = R.l_lo + R.n_levels - 1;
        for (int64_t r = t * wpb + (threadIdx.x >> 6); r < a.total_rows; r += a.G * wpb) {
            const int li = row_level(R, r);
            const int level = R.l_lo + li;
            const int64_t row = r - R.prefix[li];
            const int64_t z = row / R.x_num[li], x = row % R.x_num[li];
            const uint64_t b = a.xz_end[R.row_begin[li] + row - 1], e = a.xz_end[R.row_begin[li] + row];
            for (uint64_t i = b + lane; i < e; i += 64) {
                const int64_t y = a.y_vec[i];
                const float cur = val(a, (int64_t)i);
                float o = cur * a.f1;
                for (int k = 0; k < 2; ++k) {
                    const int sgn = k == 0 ? 1 : -1;
                    const int dy = a.dim == 0 ? sgn : 0, dx = a.dim == 1 ? sgn : 0, dz = a.dim == 2 ? sgn : 0;
                    const float ft = k == 0 ? a.f2 : a.f0;
                    float sum = 0.0f, cnt = 0.0f;
                    const int64_t ny = u16c(y + dy), nx = u16c(x + dx), nz = u16c(z + dz);
                    int64_t n = find(a, level, nz, nx, ny);
                    if (n < 0 && level > l_min) n = find(a, level - 1, nz / 2, nx / 2, ny / 2);
                    if (n >= 0) {
                        sum += val(a, n);
                        cnt += 1.0f;
                    } else if (level < l_max) {
                        const int64_t cy = u16c(ny * 2 + (dy < 0)), cx = u16c(nx * 2 + (dx < 0)), cz = u16c(nz * 2 + (dz < 0));
                        for (int c = 0; c < 4; ++c) {
                            const int o0 = c & 1, o1 = c >> 1;
                            int64_t yy = cy, xx = cx, zz = cz;
                            if (a.dim == 0) { xx += o0; zz += o1; }
                            else if (a.dim == 1) { yy += o0; zz += o1; }
                            else { yy += o0; xx += o1; }
                            const int64_t m = find(a, level + 1, u16c(zz), u16c(xx), u16c(yy));
                            if (m < 0) {
                                if (c == 0) break;
                                continue;
                            }
                            sum += val(a, m);
                            cnt += 1.0f;
                        }
                    }
                    if (cnt > 0.0f) o += ft * sum / cnt;
                    else o += ft * cur;
                }
                a.out[i] = o;
            }
        }
    }
};
}  // namespace

void face_neighbour_filter(aprh_context* ctx, const aprh_apr& a, const void* in, int in_dtype, float* out, const float* filter3, int dim) {
    if (a.total == 0) return;
    Rows R = make_rows(a.y_num, a.x_num, a.z_num, a.level_xz_vec, a.l_min, a.l_max);
    const int64_t rows = R.prefix[R.n_levels];
    const int64_t G = stride_tiles(rows, 4, 16384);
    launch_tiles<FaceFilter>(ctx, {R, a.xz_end.as<uint64_t>(), a.y_vec.as<uint16_t>(), in, in_dtype == APRH_DTYPE_U16 ? 1 : 0, out,
                                   filter3[0], filter3[1], filter3[2], dim, rows, G},
                             G);
}

void gradient_magnitude_combine(aprh_context* ctx, float* g, const float* gx, const float* gz, uint64_t n) {
    if (n == 0) return;
    const int64_t G = stride_tiles((int64_t)n, 256 * 4);
    launch_tiles<GradMag>(ctx, {g, gx, gz, n, G}, G);
}

namespace {
struct LocalStdStep {
    struct Args {
        int mode;
        float* p;
        const void* q;
        uint64_t n;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "local_std_step"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        for (uint64_t i = (uint64_t)t * blockDim.x + threadIdx.x; i < a.n; i += (uint64_t)a.G * blockDim.x) {
            if (a.mode == 0) {
                a.p[i] = (float)((const uint16_t*)a.q)[i];
            } else if (a.mode == 1) {
                a.p[i] = a.p[i] * a.p[i];
            } else {
                const float m = ((const float*)a.q)[i];
                const float d = a.p[i] - m * m;
                a.p[i] = sqrtf(d < 0.0f ? 0.0f : d);
            }
        }
    }
};
}  // namespace

namespace {
/* elementwise steps of APRNumerics::richardson_lucy_tv (APRNumerics.hpp:431-525), in the
   reference's operation order:
   mode 0, div_norm_grad's normalisation: m = sqrtf(g_z^2 + g_x^2 + g_y^2) (absent axes are
           zero), each component divided by m where m > 1e-6 (compared in double);
   mode 1, the divergence sum: a = b + a;
   mode 2, the update: a = a * b / (1 - reg * c) */
struct TvStep {
    struct Args {
        int mode;
        float* a;
        float* b;   /* mode 0: g_x or nullptr; 1, 2: read only */
        float* c;   /* mode 0: g_z or nullptr; 2: read only     */
        float reg;
        uint64_t n;
        int64_t G;
    };
    static constexpr int THREADS = 256;
    static const char* name() { return "rl_tv_step"; }
    static __device__ __forceinline__ void run(const Args& a, int64_t t) {
        for (uint64_t i = (uint64_t)t * blockDim.x + threadIdx.x; i < a.n; i += (uint64_t)a.G * blockDim.x) {
            if (a.mode == 0) {
                const float y = a.a[i], x = a.b ? a.b[i] : 0.0f, z = a.c ? a.c[i] : 0.0f;
                const float m = sqrtf(z * z + x * x + y * y);
                if ((double)m > 1e-6) {
                    if (a.c) a.c[i] = z / m;
                    if (a.b) a.b[i] = x / m;
                    a.a[i] = y / m;
                }
            } else if (a.mode == 1) {
                a.a[i] = a.b[i] + a.a[i];
            } else {
                a.a[i] = a.a[i] * a.b[i] / (1.0f - a.reg * a.c[i]);
            }
        }
    }
};
}  // namespace

void rl_tv_step(aprh_context* ctx, int mode, float* p, float* q, float* r, float reg, uint64_t n) {
    if (n == 0) return;
    const int64_t G = stride_tiles((int64_t)n, 256 * 4);
    launch_tiles<TvStep>(ctx, {mode, p, q, r, reg, n, G}, G);
}

void fill_f32(aprh_context* ctx, float* p, size_t n, float v) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_f32, kGrid, kBlock, 0, ctx->stream, p, n, v);
    HIP_LAUNCH_CHECK();
}

void local_std_step(aprh_context* ctx, int mode, float* p, const void* q, uint64_t n) {
    if (n == 0) return;
    const int64_t G = stride_tiles((int64_t)n, 256 * 4);
    launch_tiles<LocalStdStep>(ctx, {mode, p, q, n, G}, G);
}

bool patch_limits(const aprh_apr& a, aprh_patch& p) {
    /* ReconPatch::check_limits (APRReconstruction.hpp:38-64): extents ceil(dim * 2^level_delta),
       which must fit an int32 */
    if (p.level_delta > 30 || p.level_delta < -30) return false;
    int64_t m[3];
    for (int d = 0; d < 3; ++d) {
        const int64_t dim = a.org[d];
        m[d] = p.level_delta >= 0 ? (dim << p.level_delta) : ((dim + ((int64_t)1 << -p.level_delta) - 1) >> -p.level_delta);
        if (m[d] > INT32_MAX) return false;
    }
    int32_t* b[3] = {&p.y_begin, &p.x_begin, &p.z_begin};
    int32_t* e[3] = {&p.y_end, &p.x_end, &p.z_end};
    for (int d = 0; d < 3; ++d) {
        *b[d] = std::max(0, *b[d]);
        *e[d] = *e[d] < 0 ? (int32_t)m[d] : (int32_t)std::min<int64_t>(m[d], *e[d]);
    }
    return p.y_begin < p.y_end && p.x_begin < p.x_end && p.z_begin < p.z_end;
}

/* APRReconstruction::reconstruct_constant / reconstruct_level over a patch (:296-394, :397-490):
   the level image of level min(l_max, l_max + level_delta), cut to the patch (and upsampled by
   2^level_delta when level_delta > 0).  parts == nullptr: the level image. */
void reconstruct_patch(aprh_context* ctx, aprh_apr& a, const void* parts, int pdt, const aprh_patch& p, void* img, int idt) {
    const int level = a.l_max + std::min(p.level_delta, 0);
    const int s = std::max(p.level_delta, 0);
    const int64_t fy = a.y_num[level], fx = a.x_num[level], fz = a.z_num[level];
    const int64_t py = p.y_end - p.y_begin, px = p.x_end - p.x_begin, pz = p.z_end - p.z_begin;
    const size_t es = idt == APRH_DTYPE_U16 ? 2 : 4;
    auto full_image = [&](void* dst) {
        if (parts) reconstruct_any(ctx, a, parts, pdt, level, dst, idt);
        else reconstruct_level(ctx, a, level, dst, idt);
    };
    if (s == 0 && py == fy && px == fx && pz == fz) {
        full_image(img);
        return;
    }
    Buf full(ctx->pool, (size_t)(fy * fx * fz) * es);
    full_image(full.ptr());
    const int64_t n = py * px * pz;
    const int64_t G = stride_tiles(n, 256 * 4);
    if (idt == APRH_DTYPE_U16)
        launch_tiles<CropPatch<uint16_t>>(ctx, {full.as<uint16_t>(), (uint16_t*)img, fy, fx, py, px, n, p.y_begin, p.x_begin, p.z_begin, s, G}, G);
    else
        launch_tiles<CropPatch<float>>(ctx, {full.as<float>(), (float*)img, fy, fx, py, px, n, p.y_begin, p.x_begin, p.z_begin, s, G}, G);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void reconstruct_any(aprh_context* ctx, aprh_apr& a, const void* parts, int pdt, int level, void* img, int idt) {
    if (pdt == 0 && idt == 0) reconstruct<uint16_t, uint16_t>(ctx, a, (const uint16_t*)parts, level, (uint16_t*)img);
    else if (pdt == 0 && idt == 1) reconstruct<uint16_t, float>(ctx, a, (const uint16_t*)parts, level, (float*)img);
    else if (pdt == 1 && idt == 0) reconstruct<float, uint16_t>(ctx, a, (const float*)parts, level, (uint16_t*)img);
    else reconstruct<float, float>(ctx, a, (const float*)parts, level, (float*)img);
}

template void tree_mean<uint16_t>(aprh_context*, aprh_apr&, const uint16_t*, float*);
template void tree_mean<float>(aprh_context*, aprh_apr&, const float*, float*);
template void convolve<uint16_t>(aprh_context*, aprh_apr&, const std::vector<HostStencil>&, const uint16_t*, float*, bool, const float*);
template void convolve<float>(aprh_context*, aprh_apr&, const std::vector<HostStencil>&, const float*, float*, bool, const float*);

/* APRNumerics::richardson_lucy (APRNumerics.hpp:549-593): per iteration the tree means of u,
   blur = u (*) flip(psf) with relative_blur = in / blur in the convolution's epilogue, the tree
   means of relative_blur, and u *= relative_blur (*) psf in the second convolution's epilogue
   with the partial sums of the relative-update norm (stopping rule, when tol > 0) */
template <typename T>
int richardson_lucy(aprh_context* ctx, aprh_apr& a, const T* in, float* u, const std::vector<HostStencil>& psf,
                    const std::vector<HostStencil>& psf_flipped, int max_iter, bool resume, double tol, double* last_rel) {
    hipStream_t st = ctx->stream;
    const size_t P = a.total;
    if (!resume && P > 0) {
        hipLaunchKernelGGL(k_fill_f32, kGrid, kBlock, 0, st, u, P, 1.0f);
        HIP_LAUNCH_CHECK();
    }
    if (P > 0) ensure_tree(ctx, a);
    const bool sep = ctx->conv_mode == APRH_CONV_SEPARABLE;
    const bool in_u16 = std::is_same<T, uint16_t>::value;
    Buf blur(ctx->pool, std::max<size_t>(P, 1) * 4), tree(ctx->pool, std::max<uint64_t>(a.t_total, 1) * 4);
    Buf part, sums(ctx->pool, 2 * sizeof(double));
    int it = 0;
    double rel = 0;
    for (it = 0; it < max_iter && P > 0; ++it) {
        tree_mean<float>(ctx, a, u, tree.as<float>());
        conv_levels<float>(ctx, a, psf_flipped, sep, u, tree.as<float>(), blur.as<float>(), true, 1, in, in_u16, nullptr, nullptr);
        tree_mean<float>(ctx, a, blur.as<float>(), tree.as<float>());
        int64_t np = 0;
        conv_levels<float>(ctx, a, psf, sep, blur.as<float>(), tree.as<float>(), u, true, 2, nullptr, false, &part, &np);
        reduce_norm_parts(ctx, part.as<double>(), np, sums.as<double>());
        double h[2];
        HIP_CHECK(hipMemcpyAsync(h, sums.ptr(), 16, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        rel = h[1] > 0 ? std::sqrt(h[0] / h[1]) : 0.0;
        if (tol > 0 && rel < tol) {
            ++it;
            break;
        }
    }
    if (last_rel) *last_rel = rel;
    return it;
}

template int richardson_lucy<uint16_t>(aprh_context*, aprh_apr&, const uint16_t*, float*, const std::vector<HostStencil>&,
                                       const std::vector<HostStencil>&, int, bool, double, double*);
template int richardson_lucy<float>(aprh_context*, aprh_apr&, const float*, float*, const std::vector<HostStencil>&,
                                    const std::vector<HostStencil>&, int, bool, double, double*);

}  // namespace aprh
