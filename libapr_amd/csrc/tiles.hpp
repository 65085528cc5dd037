/*
 * tiles.hpp -- stage kernels written once, launched over one volume or over many.
 *
 * A stage kernel is a body struct B:
 *     struct Args { ... };                      plain data for one volume's launch
 *     static constexpr int THREADS;             block size
 *     static __device__ void run(const Args&, int64_t tile);   one block's share
 * launch_tiles<B>(ctx, args, tiles) runs tiles 0 .. tiles-1 of one volume.
 *
 * A batch plan (batch.hip) switches the context into recording mode: the same
 * pipeline code then appends (B, args, tiles) to a per-volume list instead of
 * launching.  The plan merges the lists of all volumes stage by stage and
 * issues one persistent launch per stage: block b walks tiles b, b + grid,
 * b + 2 grid, ... of the concatenated tile ranges and finds the volume owning
 * a tile by binary search over the tile prefix.  Every block of such a launch
 * iterates over whole tiles, so __syncthreads inside a body stays uniform.
 */
#ifndef APRH_TILES_HPP
#define APRH_TILES_HPP

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <typeinfo>
#include <vector>

#include "aprh_internal.hpp"

namespace aprh {

/* one way to run a stage of a volume: a body B with its arguments and tiles */
struct TileAlt {
    const std::type_info* body = nullptr;   /* identifies B                      */
    std::vector<uint8_t> args;              /* bytes of B::Args                  */
    int64_t tiles = 0;
    int threads = 0;
    /* persistent launch of B over nv volumes (args and prefix in device memory) */
    void (*multi)(hipStream_t, const void* d_args, const int64_t* d_prefix, int nv, unsigned grid) = nullptr;
    /* launch of B over one volume */
    void (*single)(hipStream_t, const void* args, int64_t tiles) = nullptr;
    const char* name = nullptr;
    const char* shape = nullptr;            /* tile shape of this alternative (or nullptr) */
};

/* one recorded launch of a volume: alts[0], or for a stage with tile-shape alternatives every
   alternative (same results, in a fixed order per stage), which a batch plan times */
struct TileRec {
    const std::type_info* body = nullptr;   /* merge key: the body of alts[0]    */
    std::vector<TileAlt> alts;
    int prefer = 0;                         /* the alternative a single volume takes */
};

struct TileRecorder {
    std::vector<TileRec> recs;
};

/* index v with prefix[v] <= t < prefix[v + 1] (prefix[0] = 0, prefix[nv] = total) */
__device__ __forceinline__ int tile_owner(const int64_t* prefix, int nv, int64_t t) {
    int lo = 0, hi = nv;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* Blocks are handed to the 8 XCDs round-robin (block b runs on XCD b % 8), and each XCD has
   its own L2.  Renumbering so that XCD k owns one contiguous run of the grid's block indices
   puts neighbouring tiles (which share halo rows) in the same L2. */
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t grid) {
    constexpr int64_t NX = 8;
    const int64_t q = grid / NX, r = grid % NX, x = b % NX;
    return x * q + min(x, r) + b / NX;
}

/* waves per SIMD a body asks the register allocator to leave room for (B::MIN_WAVES, default 1) */
template <typename B, typename = void>
struct min_waves {
    static constexpr int value = 1;
};
template <typename B>
struct min_waves<B, std::void_t<decltype(B::MIN_WAVES)>> {
    static constexpr int value = B::MIN_WAVES;
};

template <typename B>
__global__ __launch_bounds__(B::THREADS) __attribute__((amdgpu_waves_per_eu(min_waves<B>::value))) void k_tiles(typename B::Args a,
                                                                                                             int64_t tiles) {
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);
    for (int64_t t = b; t < tiles; t += gridDim.x) B::run(a, t);
}

template <typename B>
__global__ __launch_bounds__(B::THREADS) __attribute__((amdgpu_waves_per_eu(min_waves<B>::value))) void k_tiles_multi(const typename B::Args* __restrict__ a, const int64_t* __restrict__ prefix,
                                                            int nv) {
    const int64_t total = prefix[nv];
    for (int64_t t = xcd_block(blockIdx.x, gridDim.x); t < total; t += gridDim.x) {
        const int v = tile_owner(prefix, nv, t);
        B::run(a[v], t - prefix[v]);
    }
}

template <typename B>
void launch_tiles_multi(hipStream_t st, const void* d_args, const int64_t* d_prefix, int nv, unsigned grid) {
    hipLaunchKernelGGL(k_tiles_multi<B>, dim3(grid), dim3(B::THREADS), 0, st, (const typename B::Args*)d_args, d_prefix, nv);
    HIP_LAUNCH_CHECK();
}

/* blocks of a one-volume launch: one per tile up to the per-dimension work-item limit */
inline unsigned single_grid(int64_t tiles, int threads) {
    const int64_t cap = (int64_t)(0xFFFFFFFFu / (unsigned)threads);
    return (unsigned)std::min<int64_t>(tiles, cap);
}

template <typename B>
void launch_tiles_single(hipStream_t st, const void* args, int64_t tiles) {
    typename B::Args a;
    std::memcpy(&a, args, sizeof(a));
    hipLaunchKernelGGL(k_tiles<B>, dim3(single_grid(tiles, B::THREADS)), dim3(B::THREADS), 0, st, a, tiles);
    HIP_LAUNCH_CHECK();
}

/* an alternative for launch_tiles_alts: body B over `tiles` tiles, shape named for reports */
template <typename B>
TileAlt tile_alt(const typename B::Args& a, int64_t tiles, const char* shape = nullptr) {
    static_assert(std::is_trivially_copyable<typename B::Args>::value, "tile arguments must be plain data");
    TileAlt t;
    t.body = &typeid(B);
    t.args.resize(sizeof(typename B::Args));
    std::memcpy(t.args.data(), &a, sizeof(typename B::Args));
    t.tiles = tiles;
    t.threads = B::THREADS;
    t.multi = &launch_tiles_multi<B>;
    t.single = &launch_tiles_single<B>;
    t.name = B::name();
    t.shape = shape;
    return t;
}

/* a stage with tile-shape alternatives that compute the same result: one volume runs
   alts[prefer]; a batch plan records them all and keeps the fastest */
inline void launch_tiles_alts(aprh_context* ctx, std::vector<TileAlt> alts, int prefer) {
    if (alts.empty() || alts[(size_t)prefer].tiles <= 0) return;
    if (ctx->recorder) {
        TileRec r;
        r.body = alts[0].body;
        r.alts = std::move(alts);
        r.prefer = prefer;
        ctx->recorder->recs.push_back(std::move(r));
        return;
    }
    const TileAlt& t = alts[(size_t)prefer];
    t.single(ctx->stream, t.args.data(), t.tiles);
}

template <typename B>
void launch_tiles(aprh_context* ctx, const typename B::Args& a, int64_t tiles) {
    static_assert(std::is_trivially_copyable<typename B::Args>::value, "tile arguments must be plain data");
    if (tiles <= 0) return;
    if (ctx->recorder) {
        TileRec r;
        r.body = &typeid(B);
        r.alts.push_back(tile_alt<B>(a, tiles));
        ctx->recorder->recs.push_back(std::move(r));
        return;
    }
    hipLaunchKernelGGL(k_tiles<B>, dim3(single_grid(tiles, B::THREADS)), dim3(B::THREADS), 0, ctx->stream, a, tiles);
    HIP_LAUNCH_CHECK();
}

/* virtual grid of a grid-stride body: min(cap, blocks needed for n items at `per_block` each) */
inline int64_t stride_tiles(int64_t n, int64_t per_block, int64_t cap = 4096) {
    return std::max<int64_t>(1, std::min<int64_t>(cap, cdiv(n, per_block)));
}

}  // namespace aprh

#endif
