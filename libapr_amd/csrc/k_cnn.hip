/*
 * k_cnn.hip -- the learned Richardson-Lucy warm start: a 3-D convolutional
 * network (no counterpart in the reference) trained and run on the GPU with
 * hand-written MFMA kernels.
 *
 * Network: four "valid" 3x3x3 convolutions, 1 -> 16 -> 16 -> 16 -> 1
 * channels, ReLU after the first three, residual output y = x + conv4.  A
 * tile of side S + 8 gives an output tile of side S (S = 64 in use).
 * Parameters (f32 master copy, 14737 values): W1[27][16] b1[16]
 * W2[27][16][16] b2[16] W3[27][16][16] b3[16] W4[27][16] b4, tap =
 * (dz * 3 + dx) * 3 + dy.
 *
 * Precision: layers 1-3 run on v_mfma_f32_16x16x32_bf16 (bf16 weights and
 * activations, f32 accumulation); layer 4 (one output channel) runs on the
 * same MFMA with its f32 weights split into two bf16 parts, and its input
 * gradient on the VALU in f32.  Activations are stored channels-last
 * (16 bf16 = 32 B per voxel) for the convolutions and, in training, also
 * channel-planar for the weight-gradient GEMMs, whose reduction dimension is
 * the voxels.
 *
 * MFMA operand layout (16x16x32 bf16, verified on the device): lane l holds
 * A[l % 16][(l / 16) * 8 + j] and B[(l / 16) * 8 + j][l % 16] for j < 8, and
 * the result D[(l / 16) * 4 + i][l % 16] for i < 4.
 *  - convolution (forward and input gradient): M = 16 output voxels, N = 16
 *    output channels, K = pairs of taps x 16 input channels (14 MFMAs);
 *  - weight gradient: M = 16 input channels (or taps), N = 16 output
 *    channels (or taps), K = 32 voxels along y.
 */
#include <algorithm>
#include <cmath>
#include <random>

#include "aprh_internal.hpp"
#include "tiles.hpp"

namespace aprh {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int CH = 16, NT = 27, NPAIR = 14;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NT * CH, OFF_W2 = OFF_B1 + CH, OFF_B2 = OFF_W2 + NT * CH * CH, OFF_W3 = OFF_B2 + CH,
              OFF_B3 = OFF_W3 + NT * CH * CH, OFF_W4 = OFF_B3 + CH, OFF_B4 = OFF_W4 + NT * CH, N_PARAMS = OFF_B4 + 1;

/* float -> bf16, round to nearest even */
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
/* the same rounding (to nearest even) in one instruction, for finite values */
__device__ __forceinline__ uint16_t f2bf_hw(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
__device__ __forceinline__ uint32_t pack2(uint16_t lo, uint16_t hi) { return (uint32_t)lo | ((uint32_t)hi << 16); }
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
/* two floats -> two bf16 (round to nearest even, finite values) in one word */
__device__ __forceinline__ uint32_t pack2_bf16(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}
#define LDS_PTR(p) ((__attribute__((address_space(3))) void*)(p))

__device__ __forceinline__ f32x4 mfma(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

__device__ __forceinline__ void tap_offsets(int tap, int& dz, int& dx, int& dy) {
    dz = tap / 9;
    dx = (tap / 3) % 3;
    dy = tap % 3;
}

__device__ __forceinline__ void decompose(int v, int s, int& z, int& x, int& y) {
    z = v / (s * s);
    const int r = v - z * s * s;
    x = r / s;
    y = r - x * s;
}

/* ---------------- forward ---------------- */

/* The forward layers run over bricks of 16 (y) x 8 (x) x 4 (z) output voxels.  A block stages the
   brick's input, (brick + 2)^3 voxels, in LDS once, and its four waves take
   the brick's 32 y rows of 16 voxels in turn: row (kx, kz) is one 16 x 16 x 32 MFMA chain with the
   row's 16 voxels as M.  Every brick has the full layout, so a voxel's place is a shift and a mask;
   rows past the output's x or z edge are skipped, lanes past its y edge compute rows that are never
   stored (an MFMA row depends on its own A row only).  The one tap that does not exist (27, in the
   odd half of pair 13) reads tap 26's data against zero weights. */
constexpr int BY = 16, BX = 8, BZ = 4, IY = BY + 2, IX = BX + 2, IZ = BZ + 2, ROWS = BX * BZ;

/* brick br of a launch over T tiles of nby x nbx x nbz bricks, y fastest.  The kernels number
   their blocks XCD by XCD (xcd_block), so neighbouring bricks, which share their halo voxels,
   meet in one L2 */
__device__ __forceinline__ void brick_origin(uint32_t br, uint32_t nby, uint32_t nbx, uint32_t nbz, int& t, int& z0, int& x0, int& y0) {
    const uint32_t by = br % nby;
    br /= nby;
    const uint32_t bx = br % nbx;
    br /= nbx;
    const uint32_t bz = br % nbz;
    t = (int)(br / nbz);
    y0 = (int)by * BY;
    x0 = (int)bx * BX;
    z0 = (int)bz * BZ;
}

/* LDS offset (in voxels of the staged brick) of tap t */
__host__ __device__ constexpr int tap_vox(int t) { return ((t / 9) * IX + (t / 3) % 3) * IY + t % 3; }
__device__ __forceinline__ int tap_lds(int tap) { return tap_vox(min(tap, NT - 1)); }
/* odd tap 2p + 1 minus even tap 2p in voxels: 1, IY - 2 or IX * IY - 2 IY - 2; 0 for pair 13,
   whose odd half (tap 27) re-reads tap 26 */
__host__ __device__ constexpr int pair_step(int p) { return p == NPAIR - 1 ? 0 : tap_vox(2 * p + 1) - tap_vox(2 * p); }

/* one y row of a 16 -> 16 layer: lane (g, m) reads channels (g & 1) * 8 .. + 7 of tap
   2p + (g >> 1) of voxel m as the A operand of pair p.  lb[k] is the lane's offset in the brick
   plus its odd-tap step of class k (steps 0, 1, IY - 2, IX * IY - 2 IY - 2), so every read is a
   per-row base and an immediate offset */
__device__ __forceinline__ f32x4 row_mfma(const uint16_t* tile, int R, const int (&lb)[4], const u32x4 (&w)[NPAIR]) {
    const uint16_t* b0 = tile + R + lb[0];
    const uint16_t* b1 = tile + R + lb[1];
    const uint16_t* b2 = tile + R + lb[2];
    const uint16_t* b3 = tile + R + lb[3];
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int p = 0; p < NPAIR; ++p) {
        const int st = pair_step(p);
        const uint16_t* b = st == 0 ? b0 : st == 1 ? b1 : st == IY - 2 ? b2 : b3;
        acc = mfma(*(const u32x4*)(b + tap_vox(2 * p) * CH), w[p], acc);
    }
    return acc;
}
__device__ __forceinline__ void row_lane_offsets(int m, int g, int (&lb)[4]) {
    const int lo = m * CH + (g & 1) * 8, th = g >> 1;
    lb[0] = lo;
    lb[1] = lo + th * CH;
    lb[2] = lo + th * (IY - 2) * CH;
    lb[3] = lo + th * (IX * IY - 2 * IY - 2) * CH;
}
static_assert(pair_step(0) == 1 && pair_step(1) == IY - 2 && pair_step(4) == IX * IY - 2 * IY - 2, "tap steps");

/* stages the brick's input, IZ x IX x IY voxels of 16 channels from (z0, x0, y0) + off of a tile
   of side sin, as tile[(lz * IX + lx) * IY + ly][16], 16-byte chunk c at tile[8 c].
   DIRECT (the forward layers, off = 0): every chunk goes straight from global memory to LDS
   (global_load_lds_dwordx4: lane l of a wave writes the wave's base + 16 l, which is the chunk
   order), so all of a brick's loads are in flight at once and hold no registers.  Positions past
   the input's far edges load a clamped position instead of zero: in a valid convolution they feed
   only outputs past the output's edge, which are never stored.
   Otherwise (the input gradients, off = -2, zero outside the input) through registers. */
constexpr int NCH = IZ * IX * IY * 2, NCH_PT = (NCH + 255) / 256;
template <bool DIRECT>
__device__ __forceinline__ void stage_brick(uint16_t* tile, const uint16_t* __restrict__ it, int sin, int off, int z0, int x0, int y0) {
    __syncthreads();   /* the previous brick's waves are done with the tile */
    if constexpr (DIRECT) {
        const int wbase = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * 64;
#pragma unroll
        for (int j = 0; j < NCH_PT; ++j) {
            const int c = threadIdx.x + 256 * j;
            if (c < NCH) {
                const int q = c >> 1, h = c & 1;
                const int ly = q % IY, r = q / IY, lx = r % IX, lz = r / IX;
                const int zi = min(z0 + lz, sin - 1), xi = min(x0 + lx, sin - 1), yi = min(y0 + ly, sin - 1);
                __builtin_amdgcn_global_load_lds((const void*)(it + (int64_t)((zi * sin + xi) * sin + yi) * CH + h * 8),
                                                 LDS_PTR(tile + (wbase + 256 * j) * 8), 16, 0, 0);
            }
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    } else {
        for (int c = threadIdx.x; c < NCH; c += 256) {
            const int q = c >> 1, h = c & 1;
            const int ly = q % IY, r = q / IY, lx = r % IX, lz = r / IX;
            const int zi = z0 + off + lz, xi = x0 + off + lx, yi = y0 + off + ly;
            u32x4 v = {0u, 0u, 0u, 0u};
            if ((unsigned)zi < (unsigned)sin && (unsigned)xi < (unsigned)sin && (unsigned)yi < (unsigned)sin)
                v = *(const u32x4*)(it + (int64_t)((zi * sin + xi) * sin + yi) * CH + h * 8);
            *(u32x4*)&tile[c * 8] = v;
        }
    }
    __syncthreads();
}

/* layer 1: x (f32, side s0) -> a (bf16, side s1 = s0 - 2), ReLU(conv + b1).  The brick's input goes
   straight to LDS as f32 (global_load_lds_dword, clamped at the far edges as in stage_brick) and
   is rounded to bf16 when the A operand is built; lane (g, m) holds taps g * 8 .. g * 8 + 7 of
   voxel m */
constexpr int NIN1 = IZ * IX * IY, NIN1_PT = (NIN1 + 255) / 256;
__global__ __launch_bounds__(256) void k_l1_lds(const float* __restrict__ x, int s0, int s1, const u32x4* __restrict__ W1p,
                                                const float* __restrict__ b1, uint16_t* a_cl, uint16_t* a_pl, uint32_t nbricks, uint32_t nby,
                                                uint32_t nbx, uint32_t nbz) {
    __shared__ __attribute__((aligned(16))) float tile[NIN1_PT * 256];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), m = lane & 15, g = lane >> 4;
    const u32x4 wb = W1p[lane];
    const float bias = b1[m];
    int toff[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) toff[j] = tap_lds(g * 8 + j);
    const int nv = s1 * s1 * s1, nin = s0 * s0 * s0;
    for (uint32_t br = (uint32_t)xcd_block(blockIdx.x, gridDim.x); br < nbricks; br += gridDim.x) {
        int t, z0, x0, y0;
        brick_origin(br, nby, nbx, nbz, t, z0, x0, y0);
        const float* xt = x + (int64_t)t * nin;
        __syncthreads();   /* the previous brick's waves are done with the tile */
#pragma unroll
        for (int j = 0; j < NIN1_PT; ++j) {
            const int c = threadIdx.x + 256 * j;
            if (c < NIN1) {
                const int ly = c % IY, r = c / IY, lx = r % IX, lz = r / IX;
                const int zi = min(z0 + lz, s0 - 1), xi = min(x0 + lx, s0 - 1), yi = min(y0 + ly, s0 - 1);
                __builtin_amdgcn_global_load_lds((const void*)(xt + (zi * s0 + xi) * s0 + yi), LDS_PTR(tile + wave * 64 + 256 * j), 4, 0, 0);
            }
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        for (int row = wave; row < ROWS; row += 4) {
            const int kx = row % BX, kz = row / BX;
            if (x0 + kx >= s1 || z0 + kz >= s1) continue;
            const float* base = &tile[(kz * IX + kx) * IY + m];
            float f[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) f[j] = base[toff[j]];
            const u32x4 av = {pack2_bf16(f[0], f[1]), pack2_bf16(f[2], f[3]), pack2_bf16(f[4], f[5]), pack2_bf16(f[6], f[7])};
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(av, wb, acc);
            const int vrow = ((z0 + kz) * s1 + x0 + kx) * s1 + y0 + g * 4;
            const int64_t vc = ((int64_t)t * nv + vrow) * CH + m;
            const int64_t vp = ((int64_t)t * CH + m) * nv + vrow;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (y0 + g * 4 + i < s1) {
                    const uint16_t hv = f2bf_hw(fmaxf(acc[i] + bias, 0.0f));
                    a_cl[vc + i * CH] = hv;
                    if (a_pl) a_pl[vp + i] = hv;
                }
            }
        }
    }
}

/* 16 -> 16 convolution: out(v) = sum_tap W[tap] in(v + tap + off) (zero outside the input), then
   + bias, ReLU, and / or the mask (ReLU derivative of a stored activation).  K of an MFMA is a pair
   of taps x 16 input channels: lane (g, m) holds channels (g & 1) * 8 .. + 7 of tap 2p + (g >> 1) */
template <bool DIRECT>
__global__ __launch_bounds__(256) void k_conv16_lds(const uint16_t* __restrict__ in_cl, int sin, int sout, int off,
                                                    const u32x4* __restrict__ Wp, const float* __restrict__ bias,
                                                    const uint16_t* __restrict__ mask_cl, uint16_t* out_cl, uint16_t* out_pl, int relu,
                                                    uint32_t nbricks, uint32_t nby, uint32_t nbx, uint32_t nbz) {
    __shared__ __attribute__((aligned(16))) uint16_t tile[IZ * IX * IY * CH];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), m = lane & 15, g = lane >> 4;
    u32x4 w[NPAIR];
#pragma unroll
    for (int p = 0; p < NPAIR; ++p) w[p] = Wp[p * 64 + lane];
    const float bv = bias ? bias[m] : 0.0f;
    int lb[4];
    row_lane_offsets(m, g, lb);
    const int nv = sout * sout * sout;
    for (uint32_t br = (uint32_t)xcd_block(blockIdx.x, gridDim.x); br < nbricks; br += gridDim.x) {
        int t, z0, x0, y0;
        brick_origin(br, nby, nbx, nbz, t, z0, x0, y0);
        stage_brick<DIRECT>(tile, in_cl + (int64_t)t * sin * sin * sin * CH, sin, off, z0, x0, y0);
        for (int row = wave; row < ROWS; row += 4) {
            const int kx = row % BX, kz = row / BX;
            if (x0 + kx >= sout || z0 + kz >= sout) continue;
            const f32x4 acc = row_mfma(tile, (kz * IX + kx) * IY * CH, lb, w);
            /* this lane's four outputs: voxels vrow + g * 4 + i, channel m */
            const int vrow = ((z0 + kz) * sout + x0 + kx) * sout + y0 + g * 4;
            const int64_t vc = ((int64_t)t * nv + vrow) * CH + m;
            const int64_t vp = ((int64_t)t * CH + m) * nv + vrow;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (y0 + g * 4 + i < sout) {
                    float val = acc[i] + bv;
                    if (relu) val = fmaxf(val, 0.0f);
                    if (mask_cl && !(bf2f(mask_cl[vc + i * CH]) > 0.0f)) val = 0.0f;
                    const uint16_t hv = f2bf_hw(val);
                    if (out_cl) out_cl[vc + i * CH] = hv;
                    if (out_pl) out_pl[vp + i] = hv;
                }
            }
        }
    }
}

/* layer 4 + residual (+ loss gradient): y = x(centre) + b4 + sum W4 a3, on the MFMA.  The weights
   keep f32 accuracy as two bf16 parts: W4p holds bf16(W4) in output columns 0-7 and
   bf16(W4 - bf16(W4)) in columns 8-15, so column 0 + column 8 is the layer's output.  Lane
   (g, m < 4) takes voxel g * 4 + m of its row, and a row's 16 outputs leave in one coalesced store.
   With a target, also the loss gradient g4 = 2 (y - tgt) gscale and per-block partial sums of
   (y - tgt)^2 and g4 */
__global__ __launch_bounds__(256) void k_l4_lds(const uint16_t* __restrict__ a3_cl, int S, const float* __restrict__ x,
                                                const u32x4* __restrict__ W4p, const float* __restrict__ b4, const float* __restrict__ tgt,
                                                float* y, float* g4, float gscale, double* part, uint32_t nbricks, uint32_t nby, uint32_t nbx,
                                                uint32_t nbz) {
    __shared__ __attribute__((aligned(16))) uint16_t tile[IZ * IX * IY * CH];
    __shared__ double red[2][256];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), m = lane & 15, g = lane >> 4;
    u32x4 w[NPAIR];
#pragma unroll
    for (int p = 0; p < NPAIR; ++p) w[p] = W4p[p * 64 + lane];
    const float bv = b4[0];
    int lb[4];
    row_lane_offsets(m, g, lb);
    const int s3 = S + 2, s0 = S + 8;
    const int nv = S * S * S, n3 = s3 * s3 * s3, n0 = s0 * s0 * s0;
    double dsum = 0, gsum = 0;
    for (uint32_t br = (uint32_t)xcd_block(blockIdx.x, gridDim.x); br < nbricks; br += gridDim.x) {
        int t, z0, x0, y0;
        brick_origin(br, nby, nbx, nbz, t, z0, x0, y0);
        stage_brick<true>(tile, a3_cl + (int64_t)t * n3 * CH, s3, 0, z0, x0, y0);
        for (int row = wave; row < ROWS; row += 4) {
            const int kx = row % BX, kz = row / BX;
            if (x0 + kx >= S || z0 + kz >= S) continue;
            /* the residual (and target) loads go out before the MFMA chain */
            const int zo = z0 + kz, xo = x0 + kx, yo = y0 + g * 4 + m;
            const bool mine = m < 4 && yo < S;
            const int64_t i = (int64_t)t * nv + (zo * S + xo) * S + yo;
            float xv = 0.0f, tv = 0.0f;
            if (mine) {
                xv = x[(int64_t)t * n0 + ((zo + 4) * s0 + xo + 4) * s0 + yo + 4];
                if (tgt) tv = tgt[i];
            }
            const f32x4 acc = row_mfma(tile, (kz * IX + kx) * IY * CH, lb, w);
            float sum[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) sum[k] = acc[k] + __shfl_down(acc[k], 8, 16);
            if (mine) {
                const float a = m == 0 ? sum[0] : m == 1 ? sum[1] : m == 2 ? sum[2] : sum[3];
                const float out = xv + bv + a;
                y[i] = out;
                if (tgt) {
                    const float d = out - tv;
                    const float gv = 2.0f * d * gscale;
                    g4[i] = gv;
                    dsum += (double)d * (double)d;
                    gsum += (double)gv;
                }
            }
        }
    }
    if (!part) return;
    red[0][threadIdx.x] = dsum;
    red[1][threadIdx.x] = gsum;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[0][threadIdx.x] += red[0][threadIdx.x + o];
            red[1][threadIdx.x] += red[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = red[0][0];
        part[2 * blockIdx.x + 1] = red[1][0];
    }
}

/* ---------------- backward ---------------- */

/* g3(u)[c] = mask(a3(u)[c]) * sum_tap W4[tap][c] g4(u - tap); one thread per voxel of side S + 2 */
__global__ __launch_bounds__(256) void k_l4_bwd(const float* __restrict__ g4, int S, const float* __restrict__ W4,
                                                const uint16_t* __restrict__ a3_cl, uint16_t* g3_cl, uint16_t* g3_pl, int T) {
    __shared__ float w[NT * CH];
    for (int i = threadIdx.x; i < NT * CH; i += blockDim.x) w[i] = W4[i];
    __syncthreads();
    const int s3 = S + 2;
    const int nv = S * S * S, n3 = s3 * s3 * s3;
    const int64_t total = (int64_t)T * n3;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(i / n3), u = (int)(i % n3);
        int zu, xu, yu;
        decompose(u, s3, zu, xu, yu);
        float acc[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] = 0.0f;
        const float* gt = g4 + (int64_t)t * nv;
        for (int tap = 0; tap < NT; ++tap) {
            int dz, dx, dy;
            tap_offsets(tap, dz, dx, dy);
            const int z = zu - dz, x = xu - dx, y = yu - dy;
            if (z < 0 || z >= S || x < 0 || x >= S || y < 0 || y >= S) continue;
            const float gv = gt[(z * S + x) * S + y];
#pragma unroll
            for (int c = 0; c < CH; ++c) acc[c] += w[tap * CH + c] * gv;
        }
        const uint16_t* ap = a3_cl + i * CH;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const uint16_t hv = f2bf(bf2f(ap[c]) > 0.0f ? acc[c] : 0.0f);
            g3_cl[i * CH + c] = hv;
            g3_pl[((int64_t)t * CH + c) * n3 + u] = hv;
        }
    }
}

/* 8 consecutive planar bf16 values starting at y0 of a row (zero from y >= ylim) */
__device__ __forceinline__ u32x4 load8(const uint16_t* row, int y0, int ylim) {
    uint16_t h[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = (y0 + j < ylim) ? row[y0 + j] : (uint16_t)0;
    const u32x4 r = {pack2(h[0], h[1]), pack2(h[2], h[3]), pack2(h[4], h[5]), pack2(h[6], h[7])};
    return r;
}

/* dW[tap][cin][cout] = sum_v a(v + tap)[cin] g(v)[cout], db[cout] = sum_v g(v)[cout];
   one partial per wave: part[w][27 * 256 + 16] */
__global__ __launch_bounds__(256) void k_dw16(const uint16_t* __restrict__ a_pl, const uint16_t* __restrict__ g_pl, int sg, float* part,
                                              int T) {
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int sa = sg + 2;
    const int ng = sg * sg * sg, na = sa * sa * sa;
    const int nch = (sg + 31) / 32;
    const int64_t items = (int64_t)T * sg * sg * nch;
    f32x4 acc[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) acc[k] = (f32x4){0.f, 0.f, 0.f, 0.f};
    float dbs = 0.0f;
    for (int64_t it = wid; it < items; it += nw) {
        const int ch = (int)(it % nch);
        const int64_t row = it / nch;
        const int t = (int)(row / (sg * sg)), zx = (int)(row % (sg * sg));
        const int z = zx / sg, x = zx % sg;
        const int y0 = ch * 32 + g * 8;
        const u32x4 bv = load8(g_pl + ((int64_t)t * CH + m) * ng + (z * sg + x) * sg, y0, sg);
#pragma unroll
        for (int q = 0; q < 4; ++q) dbs += bf2f((uint16_t)(bv[q] & 0xffffu)) + bf2f((uint16_t)(bv[q] >> 16));
        const uint16_t* ab = a_pl + ((int64_t)t * CH + m) * na;
#pragma unroll
        for (int tap = 0; tap < NT; ++tap) {
            const int dz = tap / 9, dx = (tap / 3) % 3, dy = tap % 3;
            const u32x4 av = load8(ab + ((z + dz) * sa + x + dx) * sa + dy, y0, sg);
            acc[tap] = mfma(av, bv, acc[tap]);
        }
    }
    float* pw = part + wid * (NT * 256 + CH);
#pragma unroll
    for (int tap = 0; tap < NT; ++tap)
#pragma unroll
        for (int i = 0; i < 4; ++i) pw[tap * 256 + (g * 4 + i) * CH + m] = acc[tap][i];
    dbs += __shfl_xor(dbs, 16);
    dbs += __shfl_xor(dbs, 32);
    if (lane < 16) pw[NT * 256 + m] = dbs;
}

/* dW4[tap][cin] = sum_u a3(u)[cin] g4(u - tap): M = cin, N = tap (two blocks of 16), K = voxels u */
__global__ __launch_bounds__(256) void k_dw_l4(const uint16_t* __restrict__ a3_pl, const float* __restrict__ g4, int S, float* part, int T) {
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int s3 = S + 2;
    const int n3 = s3 * s3 * s3, nv = S * S * S;
    const int nch = (s3 + 31) / 32;
    const int64_t items = (int64_t)T * s3 * s3 * nch;
    f32x4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    for (int64_t it = wid; it < items; it += nw) {
        const int ch = (int)(it % nch);
        const int64_t row = it / nch;
        const int t = (int)(row / (s3 * s3)), zx = (int)(row % (s3 * s3));
        const int z = zx / s3, x = zx % s3;
        const int y0 = ch * 32 + g * 8;
        const u32x4 av = load8(a3_pl + ((int64_t)t * CH + m) * n3 + (z * s3 + x) * s3, y0, s3);
        const float* gt = g4 + (int64_t)t * nv;
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const int tap = nb * 16 + m;
            uint16_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (tap < NT) {
                const int dz = tap / 9, dx = (tap / 3) % 3, dy = tap % 3;
                const int vz = z - dz, vx = x - dx;
                if (vz >= 0 && vz < S && vx >= 0 && vx < S) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int vy = y0 + j - dy;
                        if (y0 + j < s3 && vy >= 0 && vy < S) h[j] = f2bf(gt[(vz * S + vx) * S + vy]);
                    }
                }
            }
            const u32x4 bv = {pack2(h[0], h[1]), pack2(h[2], h[3]), pack2(h[4], h[5]), pack2(h[6], h[7])};
            acc[nb] = mfma(av, bv, acc[nb]);
        }
    }
    float* pw = part + wid * 512;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) pw[nb * 256 + (g * 4 + i) * CH + m] = acc[nb][i];
}

/* dW1[tap][cout] = sum_v x(v + tap) g1(v)[cout], db1: M = tap (two blocks of 16), N = cout, K = voxels v */
__global__ __launch_bounds__(256) void k_dw_l1(const float* __restrict__ x, const uint16_t* __restrict__ g1_pl, int S, float* part, int T) {
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int s1 = S + 6, s0 = S + 8;
    const int n1 = s1 * s1 * s1, n0 = s0 * s0 * s0;
    const int nch = (s1 + 31) / 32;
    const int64_t items = (int64_t)T * s1 * s1 * nch;
    f32x4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    float dbs = 0.0f;
    for (int64_t it = wid; it < items; it += nw) {
        const int ch = (int)(it % nch);
        const int64_t row = it / nch;
        const int t = (int)(row / (s1 * s1)), zx = (int)(row % (s1 * s1));
        const int z = zx / s1, x0 = zx % s1;
        const int y0 = ch * 32 + g * 8;
        const u32x4 bv = load8(g1_pl + ((int64_t)t * CH + m) * n1 + (z * s1 + x0) * s1, y0, s1);
#pragma unroll
        for (int q = 0; q < 4; ++q) dbs += bf2f((uint16_t)(bv[q] & 0xffffu)) + bf2f((uint16_t)(bv[q] >> 16));
        const float* xt = x + (int64_t)t * n0;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
            const int tap = mb * 16 + m;
            uint16_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (tap < NT) {
                const int dz = tap / 9, dx = (tap / 3) % 3, dy = tap % 3;
                const float* xr = xt + ((z + dz) * s0 + x0 + dx) * s0 + dy;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (y0 + j < s1) h[j] = f2bf(xr[y0 + j]);
            }
            const u32x4 av = {pack2(h[0], h[1]), pack2(h[2], h[3]), pack2(h[4], h[5]), pack2(h[6], h[7])};
            acc[mb] = mfma(av, bv, acc[mb]);
        }
    }
    float* pw = part + wid * (512 + CH);
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) pw[mb * 256 + (g * 4 + i) * CH + m] = acc[mb][i];
    dbs += __shfl_xor(dbs, 16);
    dbs += __shfl_xor(dbs, 32);
    if (lane < 16) pw[512 + m] = dbs;
}

/* sum of per-wave partials in wave order into the gradient vector:
   mode 0: grad[off + i] = sum_w part[w * stride + i] for i < count
   mode 1 (layer 4): grad[OFF_W4 + tap * 16 + cin] = sum_w part[w * stride + nb * 256 + cin * 16 + n], tap = nb * 16 + n */
__global__ void k_reduce_parts(const float* __restrict__ part, int64_t nwaves, int stride, int count, float* grad, int off, int mode) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int src = i;
    if (mode == 1) {
        const int tap = i / CH, cin = i % CH;
        src = (tap / 16) * 256 + cin * CH + (tap % 16);
    }
    float s = 0.0f;
    for (int64_t w = 0; w < nwaves; ++w) s += part[w * stride + src];
    grad[off + i] = s;
}

__global__ void k_set_scalar(float* dst, const double* part, int nblocks, int which, double scale) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0;
    for (int i = 0; i < nblocks; ++i) s += part[2 * i + which];
    dst[0] = (float)(s * scale);
}

__global__ void k_adam(float* p, const float* __restrict__ g, float* m, float* v, int n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

/* lane-packed bf16 weights for the MFMA kernels */
__global__ void k_pack(const float* __restrict__ P, u32x4* W1p, u32x4* W2p, u32x4* W3p, u32x4* W2t, u32x4* W3t, u32x4* W4p) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;   /* 0 .. 64 + 5 * 14 * 64 */
    if (idx >= 64 + 5 * NPAIR * 64) return;
    uint16_t h[8];
    u32x4* dst;
    if (idx < 64) {
        const int l = idx, n = l & 15, g = l >> 4;
        for (int j = 0; j < 8; ++j) {
            const int k = g * 8 + j;
            h[j] = k < NT ? f2bf(P[OFF_W1 + k * CH + n]) : (uint16_t)0;
        }
        dst = W1p + l;
    } else if (idx >= 64 + 4 * NPAIR * 64) {
        /* layer 4: bf16(W4[tap][cin]) in output columns 0-7, the bf16 remainder in columns 8-15 */
        const int q = idx - 64 - 4 * NPAIR * 64;
        const int p = q / 64, l = q % 64, n = l & 15, g = l >> 4;
        const int tap = 2 * p + (g >> 1), c0 = (g & 1) * 8;
        for (int j = 0; j < 8; ++j) {
            const float wv = tap < NT ? P[OFF_W4 + tap * CH + c0 + j] : 0.0f;
            const uint16_t hi = f2bf(wv);
            h[j] = n < 8 ? hi : f2bf(wv - bf2f(hi));
        }
        dst = W4p + q;
    } else {
        const int r = idx - 64;
        const int which = r / (NPAIR * 64), q = r % (NPAIR * 64);
        const int p = q / 64, l = q % 64, n = l & 15, g = l >> 4;
        const int tap = 2 * p + (g >> 1), c0 = (g & 1) * 8;
        const int offw = (which == 0 || which == 2) ? OFF_W2 : OFF_W3;
        const bool transposed = which >= 2;
        for (int j = 0; j < 8; ++j) {
            const int cin = c0 + j;
            float wv = 0.0f;
            if (tap < NT) {
                /* forward: W[tap][cin][n]; input gradient: W[26 - tap][n][cin] */
                wv = transposed ? P[offw + ((NT - 1 - tap) * CH + n) * CH + cin] : P[offw + (tap * CH + cin) * CH + n];
            }
            h[j] = f2bf(wv);
        }
        dst = (which == 0 ? W2p : which == 1 ? W3p : which == 2 ? W2t : W3t) + p * 64 + l;
    }
    *dst = (u32x4){pack2(h[0], h[1]), pack2(h[2], h[3]), pack2(h[4], h[5]), pack2(h[6], h[7])};
}

__device__ __forceinline__ int64_t refl_any(int64_t q, int64_t n) {
    /* reflection about the edge samples, repeated for windows wider than the volume */
    if (n == 1) return 0;
    const int64_t per = 2 * (n - 1);
    q %= per;
    if (q < 0) q += per;
    return q < n ? q : per - q;
}

/* tiles of side s = size + 2 * halo starting at (origin - halo), reflect-padded, divided by scale */
template <typename T>
__global__ void k_extract(const T* __restrict__ img, int64_t yn, int64_t xn, int64_t zn, const int* __restrict__ org, int ntiles, int s, int halo,
                          float inv_scale, float* out) {
    const int64_t nv = (int64_t)s * s * s;
    const int64_t total = nv * ntiles;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(i / nv);
        const int v = (int)(i % nv);
        int z, x, y;
        decompose(v, s, z, x, y);
        const int64_t gy = refl_any((int64_t)org[3 * t] - halo + y, yn);
        const int64_t gx = refl_any((int64_t)org[3 * t + 1] - halo + x, xn);
        const int64_t gz = refl_any((int64_t)org[3 * t + 2] - halo + z, zn);
        out[i] = (float)img[(gz * xn + gx) * yn + gy] * inv_scale;
    }
}

/* tile outputs back into the image (voxels inside the volume), times scale, floored at `floor_v` */
__global__ void k_scatter_tiles(const float* __restrict__ yt, int S, const int* __restrict__ org, int ntiles, int64_t yn, int64_t xn,
                                int64_t zn, float scale, float floor_v, float* img) {
    const int64_t nv = (int64_t)S * S * S;
    const int64_t total = nv * ntiles;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(i / nv);
        int z, x, y;
        decompose((int)(i % nv), S, z, x, y);
        const int64_t gy = org[3 * t] + y, gx = org[3 * t + 1] + x, gz = org[3 * t + 2] + z;
        if (gy < yn && gx < xn && gz < zn) img[(gz * xn + gx) * yn + gy] = fmaxf(yt[i] * scale, floor_v);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sum(const T* __restrict__ v, size_t n, double* part) {
    __shared__ double red[256];
    double s = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) s += (double)v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

const dim3 kBlk(256);

}  // namespace

}  // namespace aprh

/* ======================================================================== */

struct aprh_cnn {
    aprh_context* ctx = nullptr;
    aprh::Buf P, M, V, G;                        /* params, Adam moments, gradient (f32) */
    aprh::Buf W1p, W2p, W3p, W2t, W3t, W4p;      /* packed bf16 weights                  */
    int step = 0;
    float last_loss = 0;
    /* Gaussian applied to the warm-start particles (sizes 1: none) */
    float post_sigma[3] = {0.0f, 0.0f, 0.0f};
    int32_t post_size[3] = {1, 1, 1};
};

namespace aprh {

namespace {

constexpr int WAVES = 2048;   /* waves of the weight-gradient kernels (one partial each) */

/* the bricks of T tiles of output side s, and a grid of at most cap blocks over them */
struct Bricks {
    uint32_t n, ny, nx, nz;
    unsigned grid;
};
Bricks bricks(int T, int s, int64_t cap) {
    Bricks b;
    b.ny = (uint32_t)cdiv(s, BY);
    b.nx = (uint32_t)cdiv(s, BX);
    b.nz = (uint32_t)cdiv(s, BZ);
    const int64_t n = (int64_t)T * b.ny * b.nx * b.nz;
    if (n > 0x7FFFFFFF) throw Error(APRH_ERR_INVALID_ARGUMENT, "too many tiles in one network pass");
    b.n = (uint32_t)n;
    b.grid = (unsigned)std::min<int64_t>(n, cap);
    return b;
}

/* 16 -> 16 convolution of T tiles (input side sin, output side sout, input offset off) */
void conv16(hipStream_t st, const uint16_t* in_cl, int sin, int sout, int off, const u32x4* Wp, const float* bias,
            const uint16_t* mask_cl, uint16_t* out_cl, uint16_t* out_pl, int relu, int T) {
    const Bricks b = bricks(T, sout, 65536);
    if (off == 0)
        hipLaunchKernelGGL(k_conv16_lds<true>, dim3(b.grid), kBlk, 0, st, in_cl, sin, sout, off, Wp, bias, mask_cl, out_cl, out_pl, relu, b.n, b.ny,
                           b.nx, b.nz);
    else
        hipLaunchKernelGGL(k_conv16_lds<false>, dim3(b.grid), kBlk, 0, st, in_cl, sin, sout, off, Wp, bias, mask_cl, out_cl, out_pl, relu, b.n,
                           b.ny, b.nx, b.nz);
    HIP_LAUNCH_CHECK();
}

void pack(aprh_cnn* n) {
    hipLaunchKernelGGL(k_pack, dim3((64 + 5 * NPAIR * 64 + 255) / 256), kBlk, 0, n->ctx->stream, n->P.as<float>(), n->W1p.as<u32x4>(),
                       n->W2p.as<u32x4>(), n->W3p.as<u32x4>(), n->W2t.as<u32x4>(), n->W3t.as<u32x4>(), n->W4p.as<u32x4>());
    HIP_LAUNCH_CHECK();
}

struct Acts {
    Buf a1c, a1p, a2c, a2p, a3c, a3p;
};

/* blocks of the layer-4 kernel (one loss partial each) */
int l4_blocks(int T, int S) { return (int)bricks(T, S, 4096).grid; }

/* forward of T tiles (input side S + 8); planar copies when training; part holds l4_blocks(T, S)
   pairs of loss partials */
void forward(aprh_cnn* n, const float* x, int T, int S, float* y, Acts& A, bool train, const float* tgt, float* g4, double* part) {
    hipStream_t st = n->ctx->stream;
    const int s0 = S + 8, s1 = S + 6, s2 = S + 4, s3 = S + 2;
    auto vox = [&](int s) { return (size_t)T * s * s * s; };
    A.a1c.reset(n->ctx->pool, vox(s1) * CH * 2);
    A.a2c.reset(n->ctx->pool, vox(s2) * CH * 2);
    A.a3c.reset(n->ctx->pool, vox(s3) * CH * 2);
    if (train) {
        A.a1p.reset(n->ctx->pool, vox(s1) * CH * 2);
        A.a2p.reset(n->ctx->pool, vox(s2) * CH * 2);
        A.a3p.reset(n->ctx->pool, vox(s3) * CH * 2);
    }
    const float* P = n->P.as<float>();
    {
        const Bricks b = bricks(T, s1, 65536);
        hipLaunchKernelGGL(k_l1_lds, dim3(b.grid), kBlk, 0, st, x, s0, s1, (const u32x4*)n->W1p.as<u32x4>(), P + OFF_B1, A.a1c.as<uint16_t>(),
                           train ? A.a1p.as<uint16_t>() : (uint16_t*)nullptr, b.n, b.ny, b.nx, b.nz);
        HIP_LAUNCH_CHECK();
    }
    conv16(st, A.a1c.as<uint16_t>(), s1, s2, 0, n->W2p.as<u32x4>(), P + OFF_B2, nullptr, A.a2c.as<uint16_t>(),
           train ? A.a2p.as<uint16_t>() : nullptr, 1, T);
    conv16(st, A.a2c.as<uint16_t>(), s2, s3, 0, n->W3p.as<u32x4>(), P + OFF_B3, nullptr, A.a3c.as<uint16_t>(),
           train ? A.a3p.as<uint16_t>() : nullptr, 1, T);
    const Bricks b4 = bricks(T, S, 4096);
    hipLaunchKernelGGL(k_l4_lds, dim3(b4.grid), kBlk, 0, st, (const uint16_t*)A.a3c.as<uint16_t>(), S, x, (const u32x4*)n->W4p.as<u32x4>(),
                       P + OFF_B4, tgt, y, g4, 1.0f / (float)vox(S), part, b4.n, b4.ny, b4.nx, b4.nz);
    HIP_LAUNCH_CHECK();
}

void train_step(aprh_cnn* n, const float* x, const float* tgt, int T, int S, float lr, float* loss_out, float* grads_out) {
    hipStream_t st = n->ctx->stream;
    Pool* pool = n->ctx->pool;
    const int s1 = S + 6, s2 = S + 4, s3 = S + 2;
    auto vox = [&](int s) { return (size_t)T * s * s * s; };
    const int nblk_l4 = l4_blocks(T, S);
    Buf y(pool, vox(S) * 4), g4(pool, vox(S) * 4), part(pool, (size_t)nblk_l4 * 2 * 8);
    Acts A;
    forward(n, x, T, S, y.as<float>(), A, true, tgt, g4.as<float>(), part.as<double>());
    float* G = n->G.as<float>();
    const float* P = n->P.as<float>();
    /* loss and db4 */
    Buf dloss(pool, 4);
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, st, dloss.as<float>(), (const double*)part.as<double>(), nblk_l4, 0,
                       1.0 / (double)vox(S));
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, st, G + OFF_B4, (const double*)part.as<double>(), nblk_l4, 1, 1.0);
    HIP_LAUNCH_CHECK();
    /* layer 4 */
    Buf g3c(pool, vox(s3) * CH * 2), g3p(pool, vox(s3) * CH * 2);
    hipLaunchKernelGGL(k_l4_bwd, dim3((unsigned)std::min<size_t>((vox(s3) + 255) / 256, 8192)), kBlk, 0, st, (const float*)g4.as<float>(), S,
                       P + OFF_W4, (const uint16_t*)A.a3c.as<uint16_t>(), g3c.as<uint16_t>(), g3p.as<uint16_t>(), T);
    HIP_LAUNCH_CHECK();
    Buf pbuf(pool, (size_t)WAVES * (NT * 256 + CH) * 4);
    const dim3 wgrid(WAVES / 4);
    hipLaunchKernelGGL(k_dw_l4, wgrid, kBlk, 0, st, (const uint16_t*)A.a3p.as<uint16_t>(), (const float*)g4.as<float>(), S,
                       pbuf.as<float>(), T);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_parts, dim3((NT * CH + 255) / 256), kBlk, 0, st, (const float*)pbuf.as<float>(), (int64_t)WAVES, 512,
                       NT * CH, G, OFF_W4, 1);
    HIP_LAUNCH_CHECK();
    /* layer 3 weights, then the gradient of a2 */
    hipLaunchKernelGGL(k_dw16, wgrid, kBlk, 0, st, (const uint16_t*)A.a2p.as<uint16_t>(), (const uint16_t*)g3p.as<uint16_t>(), s3,
                       pbuf.as<float>(), T);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_parts, dim3((NT * 256 + CH + 255) / 256), kBlk, 0, st, (const float*)pbuf.as<float>(), (int64_t)WAVES,
                       NT * 256 + CH, NT * 256 + CH, G, OFF_W3, 0);
    HIP_LAUNCH_CHECK();
    auto grid = [&](size_t v) { return dim3((unsigned)std::min<size_t>((v / 16 + 3) / 4 + 1, 16384)); };
    Buf g2c(pool, vox(s2) * CH * 2), g2p(pool, vox(s2) * CH * 2);
    conv16(st, g3c.as<uint16_t>(), s3, s2, -2, n->W3t.as<u32x4>(), nullptr, A.a2c.as<uint16_t>(), g2c.as<uint16_t>(), g2p.as<uint16_t>(),
           0, T);
    /* layer 2 weights, then the gradient of a1 */
    hipLaunchKernelGGL(k_dw16, wgrid, kBlk, 0, st, (const uint16_t*)A.a1p.as<uint16_t>(), (const uint16_t*)g2p.as<uint16_t>(), s2,
                       pbuf.as<float>(), T);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_parts, dim3((NT * 256 + CH + 255) / 256), kBlk, 0, st, (const float*)pbuf.as<float>(), (int64_t)WAVES,
                       NT * 256 + CH, NT * 256 + CH, G, OFF_W2, 0);
    HIP_LAUNCH_CHECK();
    Buf g1p(pool, vox(s1) * CH * 2);
    conv16(st, g2c.as<uint16_t>(), s2, s1, -2, n->W2t.as<u32x4>(), nullptr, A.a1c.as<uint16_t>(), nullptr, g1p.as<uint16_t>(), 0, T);
    /* layer 1 */
    hipLaunchKernelGGL(k_dw_l1, wgrid, kBlk, 0, st, x, (const uint16_t*)g1p.as<uint16_t>(), S, pbuf.as<float>(), T);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_parts, dim3((NT * CH + CH + 255) / 256), kBlk, 0, st, (const float*)pbuf.as<float>(), (int64_t)WAVES,
                       512 + CH, NT * CH, G, OFF_W1, 0);
    HIP_LAUNCH_CHECK();
    /* db1 sits after the 2 x 256 block */
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), kBlk, 0, st, (const float*)(pbuf.as<float>() + 512), (int64_t)WAVES, 512 + CH, CH, G, OFF_B1,
                       0);
    HIP_LAUNCH_CHECK();
    if (grads_out) HIP_CHECK(hipMemcpyAsync(grads_out, G, N_PARAMS * 4, hipMemcpyDeviceToHost, st));
    if (lr > 0) {
        n->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.0f - std::pow(b1, (float)n->step), bc2 = 1.0f - std::pow(b2, (float)n->step);
        hipLaunchKernelGGL(k_adam, dim3((N_PARAMS + 255) / 256), kBlk, 0, st, n->P.as<float>(), (const float*)G, n->M.as<float>(),
                           n->V.as<float>(), N_PARAMS, lr, b1, b2, 1e-8f, bc1, bc2);
        HIP_LAUNCH_CHECK();
        pack(n);
    }
    float loss = 0;
    HIP_CHECK(hipMemcpyAsync(&loss, dloss.ptr(), 4, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    n->last_loss = loss;
    if (loss_out) *loss_out = loss;
}

double device_mean(aprh_context* ctx, const void* v, int dt, size_t n) {
    const int nb = 1024;
    Buf part(ctx->pool, nb * 8);
    if (dt == APRH_DTYPE_U16)
        hipLaunchKernelGGL(k_sum<uint16_t>, dim3(nb), kBlk, 0, ctx->stream, (const uint16_t*)v, n, part.as<double>());
    else
        hipLaunchKernelGGL(k_sum<float>, dim3(nb), kBlk, 0, ctx->stream, (const float*)v, n, part.as<double>());
    HIP_LAUNCH_CHECK();
    std::vector<double> h(nb);
    HIP_CHECK(hipMemcpyAsync(h.data(), part.ptr(), nb * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double s = 0;
    for (double d : h) s += d;
    return n ? s / (double)n : 0.0;
}

/* the network over a whole volume in 64^3 tiles; out = y * scale (floored) */
void apply_volume(aprh_cnn* n, const void* img, int dt, int64_t yn, int64_t xn, int64_t zn, float scale, float floor_v, float* out) {
    hipStream_t st = n->ctx->stream;
    Pool* pool = n->ctx->pool;
    const int S = 64, s0 = S + 8;
    const int ty = (int)cdiv(yn, S), tx = (int)cdiv(xn, S), tz = (int)cdiv(zn, S);
    const int ntiles = ty * tx * tz;
    std::vector<int> org((size_t)ntiles * 3);
    for (int k = 0, z = 0; z < tz; ++z)
        for (int x = 0; x < tx; ++x)
            for (int y = 0; y < ty; ++y, ++k) {
                org[3 * k] = y * S;
                org[3 * k + 1] = x * S;
                org[3 * k + 2] = z * S;
            }
    Buf dorg(pool, org.size() * 4);
    HIP_CHECK(hipMemcpyAsync(dorg.ptr(), org.data(), org.size() * 4, hipMemcpyHostToDevice, st));
    const int B = 32;   /* tiles per batch */
    Buf xin(pool, (size_t)B * s0 * s0 * s0 * 4), y(pool, (size_t)B * S * S * S * 4);
    for (int t0 = 0; t0 < ntiles; t0 += B) {
        const int T = std::min(B, ntiles - t0);
        const int* o = dorg.as<int>() + 3 * t0;
        if (dt == APRH_DTYPE_U16)
            hipLaunchKernelGGL(k_extract<uint16_t>, dim3(8192), kBlk, 0, st, (const uint16_t*)img, yn, xn, zn, o, T, s0, 4, 1.0f / scale,
                               xin.as<float>());
        else
            hipLaunchKernelGGL(k_extract<float>, dim3(8192), kBlk, 0, st, (const float*)img, yn, xn, zn, o, T, s0, 4, 1.0f / scale,
                               xin.as<float>());
        HIP_LAUNCH_CHECK();
        Acts A;
        forward(n, xin.as<float>(), T, S, y.as<float>(), A, false, nullptr, nullptr, nullptr);
        hipLaunchKernelGGL(k_scatter_tiles, dim3(8192), kBlk, 0, st, (const float*)y.as<float>(), S, o, T, yn, xn, zn, scale, floor_v, out);
        HIP_LAUNCH_CHECK();
    }
    HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace

}  // namespace aprh

using namespace aprh;

namespace {
void check_cnn(aprh_context* ctx, aprh_cnn* n) {
    check_ctx(ctx);
    if (!n) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL network");
    if (n->ctx != ctx) throw Error(APRH_ERR_INVALID_ARGUMENT, "network belongs to another context");
}
void check_tile_size(int S) {
    if (S < 1 || S > 128) throw Error(APRH_ERR_INVALID_ARGUMENT, "tile side must be in 1..128");
}
}  // namespace

extern "C" {

int aprh_cnn_param_count(void) { return N_PARAMS; }

int aprh_cnn_create(aprh_context* ctx, uint64_t seed, aprh_cnn** out) {
    APRH_TRY
    check_ctx(ctx);
    if (!out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL out");
    aprh_cnn* n = new aprh_cnn();
    n->ctx = ctx;
    try {
        n->P.reset(ctx->pool, N_PARAMS * 4);
        n->M.reset(ctx->pool, N_PARAMS * 4);
        n->V.reset(ctx->pool, N_PARAMS * 4);
        n->G.reset(ctx->pool, N_PARAMS * 4);
        n->W1p.reset(ctx->pool, 64 * 16);
        n->W2p.reset(ctx->pool, NPAIR * 64 * 16);
        n->W3p.reset(ctx->pool, NPAIR * 64 * 16);
        n->W2t.reset(ctx->pool, NPAIR * 64 * 16);
        n->W3t.reset(ctx->pool, NPAIR * 64 * 16);
        n->W4p.reset(ctx->pool, NPAIR * 64 * 16);
        /* He initialisation for layers 1-3, small last layer (the network starts near the identity) */
        std::mt19937_64 rng(seed);
        std::vector<float> p(N_PARAMS, 0.0f);
        std::normal_distribution<float> n1(0.0f, std::sqrt(2.0f / 27.0f)), n2(0.0f, std::sqrt(2.0f / 432.0f)), n4(0.0f, 1e-3f);
        for (int i = OFF_W1; i < OFF_B1; ++i) p[i] = n1(rng);
        for (int i = OFF_W2; i < OFF_B2; ++i) p[i] = n2(rng);
        for (int i = OFF_W3; i < OFF_B3; ++i) p[i] = n2(rng);
        for (int i = OFF_W4; i < OFF_B4; ++i) p[i] = n4(rng);
        HIP_CHECK(hipMemcpyAsync(n->P.ptr(), p.data(), N_PARAMS * 4, hipMemcpyHostToDevice, ctx->stream));
        HIP_CHECK(hipMemsetAsync(n->M.ptr(), 0, N_PARAMS * 4, ctx->stream));
        HIP_CHECK(hipMemsetAsync(n->V.ptr(), 0, N_PARAMS * 4, ctx->stream));
        HIP_CHECK(hipMemsetAsync(n->G.ptr(), 0, N_PARAMS * 4, ctx->stream));
        pack(n);
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        delete n;
        throw;
    }
    *out = n;
    APRH_CATCH
}

int aprh_cnn_destroy(aprh_cnn* n) {
    APRH_TRY
    if (!n) return APRH_OK;
    (void)hipSetDevice(n->ctx->device);
    (void)hipStreamSynchronize(n->ctx->stream);
    delete n;
    APRH_CATCH
}

int aprh_cnn_get_params(aprh_context* ctx, aprh_cnn* n, float* out) {
    APRH_TRY
    check_cnn(ctx, n);
    if (!out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL out");
    copy_any(ctx, out, n->P.ptr(), N_PARAMS * 4);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    APRH_CATCH
}

int aprh_cnn_set_params(aprh_context* ctx, aprh_cnn* n, const float* in) {
    APRH_TRY
    check_cnn(ctx, n);
    if (!in) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL params");
    copy_any(ctx, n->P.ptr(), in, N_PARAMS * 4);
    HIP_CHECK(hipMemsetAsync(n->M.ptr(), 0, N_PARAMS * 4, ctx->stream));
    HIP_CHECK(hipMemsetAsync(n->V.ptr(), 0, N_PARAMS * 4, ctx->stream));
    n->step = 0;
    pack(n);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    APRH_CATCH
}

int aprh_cnn_forward_tiles(aprh_context* ctx, aprh_cnn* n, const float* in, int32_t T, int32_t S, float* out) {
    APRH_TRY
    check_cnn(ctx, n);
    check_tile_size(S);
    if (!in || !out || T < 1) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad tiles");
    const size_t ni = (size_t)T * (S + 8) * (S + 8) * (S + 8), no = (size_t)T * S * S * S;
    DevIn x(ctx, in, ni * 4);
    DevOut y(ctx, out, no * 4);
    Acts A;
    forward(n, (const float*)x.ptr, T, S, (float*)y.ptr, A, false, nullptr, nullptr, nullptr);
    y.finish(ctx);
    APRH_CATCH
}

int aprh_cnn_train_step(aprh_context* ctx, aprh_cnn* n, const float* in, const float* target, int32_t T, int32_t S, float lr, float* loss,
                        float* grads_out) {
    APRH_TRY
    check_cnn(ctx, n);
    check_tile_size(S);
    if (!in || !target || T < 1 || lr < 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad training batch");
    const size_t ni = (size_t)T * (S + 8) * (S + 8) * (S + 8), no = (size_t)T * S * S * S;
    DevIn x(ctx, in, ni * 4), t(ctx, target, no * 4);
    float* gh = nullptr;
    std::vector<float> gbuf;
    if (grads_out) {
        gbuf.resize(N_PARAMS);
        gh = gbuf.data();
    }
    train_step(n, (const float*)x.ptr, (const float*)t.ptr, T, S, lr, loss, gh);
    if (grads_out) copy_any(ctx, grads_out, gbuf.data(), N_PARAMS * 4);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    APRH_CATCH
}

int aprh_cnn_extract_tiles(aprh_context* ctx, const void* image, int dtype, int64_t y, int64_t x, int64_t z, const int32_t* origins,
                           int32_t T, int32_t size, int32_t halo, float scale, float* out) {
    APRH_TRY
    check_ctx(ctx);
    if (!image || !origins || !out || T < 1 || size < 1 || halo < 0 || !(scale > 0) || y < 1 || x < 1 || z < 1)
        throw Error(APRH_ERR_INVALID_ARGUMENT, "bad tile extraction");
    if (dtype != APRH_DTYPE_U16 && dtype != APRH_DTYPE_F32) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad dtype");
    const int s = size + 2 * halo;
    const size_t n = (size_t)(y * x * z);
    DevIn im(ctx, image, n * (dtype == APRH_DTYPE_U16 ? 2 : 4));
    DevIn org(ctx, origins, (size_t)T * 3 * 4);
    DevOut o(ctx, out, (size_t)T * s * s * s * 4);
    if (dtype == APRH_DTYPE_U16)
        hipLaunchKernelGGL(k_extract<uint16_t>, dim3(8192), kBlk, 0, ctx->stream, (const uint16_t*)im.ptr, y, x, z, (const int*)org.ptr, T, s,
                           halo, 1.0f / scale, (float*)o.ptr);
    else
        hipLaunchKernelGGL(k_extract<float>, dim3(8192), kBlk, 0, ctx->stream, (const float*)im.ptr, y, x, z, (const int*)org.ptr, T, s, halo,
                           1.0f / scale, (float*)o.ptr);
    HIP_LAUNCH_CHECK();
    o.finish(ctx);
    APRH_CATCH
}

int aprh_mean(aprh_context* ctx, const void* data, int dtype, uint64_t n, double* mean) {
    APRH_TRY
    check_ctx(ctx);
    if (!data || !mean || (dtype != APRH_DTYPE_U16 && dtype != APRH_DTYPE_F32)) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad arguments");
    DevIn d(ctx, data, std::max<uint64_t>(n, 1) * (dtype == APRH_DTYPE_U16 ? 2 : 4));
    *mean = device_mean(ctx, d.ptr, dtype, n);
    APRH_CATCH
}

int aprh_cnn_apply(aprh_context* ctx, aprh_cnn* n, const void* image, int dtype, int64_t y, int64_t x, int64_t z, float scale, float* out) {
    APRH_TRY
    check_cnn(ctx, n);
    if (!image || !out || y < 1 || x < 1 || z < 1) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad arguments");
    if (dtype != APRH_DTYPE_U16 && dtype != APRH_DTYPE_F32) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad dtype");
    const size_t nv = (size_t)(y * x * z);
    DevIn im(ctx, image, nv * (dtype == APRH_DTYPE_U16 ? 2 : 4));
    DevOut o(ctx, out, nv * 4);
    if (!(scale > 0)) scale = (float)device_mean(ctx, im.ptr, dtype, nv);
    if (!(scale > 0)) scale = 1.0f;
    apply_volume(n, im.ptr, dtype, y, x, z, scale, -INFINITY, (float*)o.ptr);
    o.finish(ctx);
    APRH_CATCH
}

int aprh_cnn_warm_start(aprh_context* ctx, aprh_cnn* n, aprh_apr* a, const void* parts, int dtype, float* u0) {
    APRH_TRY
    check_cnn(ctx, n);
    if (!a || !parts || !u0) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (dtype != APRH_DTYPE_U16 && dtype != APRH_DTYPE_F32) throw Error(APRH_ERR_INVALID_ARGUMENT, "bad dtype");
    const int64_t yn = a->org[0], xn = a->org[1], zn = a->org[2];
    const size_t nv = (size_t)(yn * xn * zn);
    DevIn p(ctx, parts, std::max<uint64_t>(a->total, 1) * (dtype == APRH_DTYPE_U16 ? 2 : 4));
    DevOut u(ctx, u0, std::max<uint64_t>(a->total, 1) * 4);
    Buf rec(ctx->pool, nv * 4), pred(ctx->pool, nv * 4);
    /* piecewise-constant image of the particles (APRReconstruction::reconstruct_constant) */
    reconstruct_any(ctx, *a, p.ptr, dtype, a->l_max, rec.ptr(), APRH_DTYPE_F32);
    float scale = (float)device_mean(ctx, rec.ptr(), APRH_DTYPE_F32, nv);
    if (!(scale > 0)) scale = 1.0f;
    /* the estimate must stay positive for the multiplicative updates */
    apply_volume(n, rec.ptr(), APRH_DTYPE_F32, yn, xn, zn, scale, 1e-3f * scale, pred.as<float>());
    launch_sample_f32(ctx, *a, pred.as<float>(), (float*)u.ptr);
    if (n->post_size[0] * n->post_size[1] * n->post_size[2] > 1 && a->total > 0) {
        /* the post filter: APRFilter::convolve_pencil of the estimate with a normalised Gaussian
           (downsampled per level, reflected boundary) */
        HostStencil hs;
        hs.sy = n->post_size[0];
        hs.sx = n->post_size[1];
        hs.sz = n->post_size[2];
        hs.w.assign((size_t)(hs.sy * hs.sx * hs.sz), 0.0f);
        if (aprh_gaussian_stencil(n->post_sigma, n->post_size, 1, hs.w.data()) != APRH_OK)
            throw Error(APRH_ERR_INVALID_ARGUMENT, "post filter stencil");
        const std::vector<HostStencil> v = downsampled_stencils(hs, a->l_max - a->l_min, true);
        Buf tmp(ctx->pool, a->total * 4);
        convolve<float>(ctx, *a, v, (const float*)u.ptr, tmp.as<float>(), true);
        launch_copy(ctx, u.ptr, tmp.ptr(), a->total * 4);
    }
    u.finish(ctx);
    APRH_CATCH
}

int aprh_cnn_set_post_filter(aprh_cnn* n, const float* sigma3, const int32_t* size3) {
    APRH_TRY
    if (!n) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL network");
    if (!sigma3 || !size3) {
        for (int d = 0; d < 3; ++d) n->post_sigma[d] = 0.0f, n->post_size[d] = 1;
    } else {
        for (int d = 0; d < 3; ++d) {
            if (size3[d] < 1 || size3[d] > 15 || size3[d] % 2 == 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "post filter sizes: odd, 1..15");
            if (size3[d] > 1 && !(sigma3[d] > 0.0f)) throw Error(APRH_ERR_INVALID_ARGUMENT, "post filter sigma must be positive");
        }
        for (int d = 0; d < 3; ++d) n->post_sigma[d] = sigma3[d], n->post_size[d] = size3[d];
    }
    APRH_CATCH
}

}  // extern "C"
