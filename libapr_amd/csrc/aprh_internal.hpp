/*
 * aprh_internal.hpp -- shared internals of libaprh (HIP, gfx950).
 *
 * Error handling: every failure throws aprh::Error carrying an aprh_status
 * code; the extern "C" layer catches it, stores the message for
 * aprh_last_error() and returns the code.
 *
 * Memory: each context owns a caching device allocator (size-bucketed free
 * lists) so repeated calls on same-sized volumes perform no hipMalloc.
 */
#ifndef APRH_INTERNAL_HPP
#define APRH_INTERNAL_HPP

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/aprh.h"

namespace aprh {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define HIP_CHECK(x)                                                   \
    do {                                                               \
        hipError_t e_ = (x);                                           \
        if (e_ != hipSuccess) ::aprh::throw_hip(e_, #x, __FILE__, __LINE__); \
    } while (0)

#define HIP_LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

/* size-bucketed caching allocator (per context / device) */
class Pool {
public:
    explicit Pool(int device) : device_(device) {}
    ~Pool();
    void* get(size_t bytes);
    void put(void* p);
    void release_all();
    size_t bytes_cached() const { return cached_; }
    /* keep mode: returned blocks stay allocated and are never handed out again (a batch
       plan records every volume's stages before any of them runs) */
    void set_keep(bool k) { keep_ = k; }

private:
    static size_t bucket(size_t b);
    int device_;
    std::multimap<size_t, void*> free_;
    std::map<void*, size_t> live_;
    size_t cached_ = 0;
    bool keep_ = false;
};

/* device buffer leased from a Pool; returned on destruction */
class Buf {
public:
    Buf() {}
    Buf(Pool* pool, size_t bytes) { reset(pool, bytes); }
    ~Buf() { release(); }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    Buf(Buf&& o) noexcept { *this = std::move(o); }
    Buf& operator=(Buf&& o) noexcept {
        if (this != &o) {
            release();
            pool_ = o.pool_; p_ = o.p_; bytes_ = o.bytes_;
            o.pool_ = nullptr; o.p_ = nullptr; o.bytes_ = 0;
        }
        return *this;
    }
    void reset(Pool* pool, size_t bytes) {
        release();
        pool_ = pool;
        bytes_ = bytes;
        p_ = bytes ? pool->get(bytes) : nullptr;
    }
    void release() {
        if (p_ && pool_) pool_->put(p_);
        p_ = nullptr; bytes_ = 0;
    }
    template <typename T> T* as() const { return (T*)p_; }
    void* ptr() const { return p_; }
    size_t bytes() const { return bytes_; }

private:
    Pool* pool_ = nullptr;
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

/* host loop over [0, n) split into contiguous ranges, one per hardware thread; f(begin, end) */
template <typename F>
void host_parallel_for(int64_t n, F f) {
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t b = n * t / nt, e = n * (t + 1) / nt;
        th.emplace_back([=] { f(b, e); });
    }
    for (auto& t : th) t.join();
}

/* timed stages of get_apr / sample_image (APRConverter's timers, split per kernel group) */
enum Stage {
    ST_OFFSET = 0,   /* min/max + bspline offset               */
    ST_SMOOTH_Y,     /* bspline_filt_rec_y (offset copy fused)  */
    ST_SMOOTH_X,     /* bspline_filt_rec_x                      */
    ST_SMOOTH_Z,     /* bspline_filt_rec_z                      */
    ST_GRADIENT,     /* gradient magnitude + max/mean ds        */
    ST_INV_BSPLINE,  /* calc_inv_bspline_y/x/z                  */
    ST_LIS,          /* local intensity scale                   */
    ST_LEVELS,       /* applyParameters + computeLevels         */
    ST_PULLING,      /* level pyramid + pulling scheme          */
    ST_STRUCTURE,    /* linear structure                        */
    ST_SAMPLE,       /* sample_image                            */
    ST_COUNT
};

}  // namespace aprh

namespace aprh { struct TileRecorder; }

namespace aprh {
/* one z block of a blocked conversion (aprh_get_apr_by_block_u16): the patch's pipeline stops
   after computeLevels (with the whole image's l_max) and seeds the global particle-cell tree */
struct BlockReq {
    int l_max = 0;                   /* the whole image's l_max                     */
    const aprh_apr* global = nullptr;
    uint8_t* pct = nullptr;          /* global particle-cell tree (pct_layout)      */
    std::vector<int64_t> off;
    int64_t z0 = 0, gl = 0, gr = 0;  /* block start, ghost planes below and above  */
};
/* conversion with sampling (aprh_get_apr_sample_u16): the row fill of the structure also
   gathers the image's particle intensities when they fit into parts */
struct SampleReq {
    const uint16_t* img = nullptr;   /* device copy of the converted image */
    uint16_t* parts = nullptr;
    uint64_t capacity = 0;
    bool done = false;
    Buf pyr;                         /* the image's mean pyramid (levels l_min .. l_max-1)     */
    std::vector<int64_t> pyr_off;    /* element offset of each level in pyr                   */
    bool top_built = false;          /* level l_max-1 already written (by the y pass)         */
};
}  // namespace aprh

struct aprh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    aprh::Pool* pool = nullptr;
    int keep_intermediates = 0;
    int timing = 0;
    hipEvent_t ev[aprh::ST_COUNT + 1] = {};
    bool ev_marked[aprh::ST_COUNT + 1] = {};
    float stage_ms[aprh::ST_COUNT] = {};
    int* d_flags = nullptr;   /* small device scratch: overflow counters etc. */
    int conv_mode = 0;        /* APRH_CONV_EXACT / APRH_CONV_SEPARABLE */
    /* asynchronous conversion (graph capture of a batch): no host synchronisation, no
       memset nodes; per-volume results left in device slots; y_vec / xz_end preallocated */
    int async_mode = 0;
    uint64_t* slot_total = nullptr;   /* particle count                        */
    float* slot_offset = nullptr;     /* B-spline offset                       */
    int* slot_err = nullptr;          /* 3 IIR overflow counters (y, x, z)     */
    /* recording mode of a batch plan: stage launches are appended here instead of run */
    aprh::TileRecorder* recorder = nullptr;
    aprh::SampleReq* sample_req = nullptr;   /* set for one aprh_get_apr_sample_u16 call */
    aprh::BlockReq* block_req = nullptr;     /* set while a blocked conversion runs a patch */
};

/* device-resident APR: GenInfo + LinearAccess (+ optional intermediates) */
struct aprh_apr {
    aprh_context* ctx = nullptr;
    int l_min = 1, l_max = 1;
    int64_t org[3] = {0, 0, 0};
    std::vector<int64_t> y_num, x_num, z_num;      /* per level, index 0..l_max */
    std::vector<uint64_t> level_xz_vec;            /* host copy (l_max+2)       */
    uint64_t total = 0;
    uint64_t n_xz = 0;                             /* size of xz_end_vec        */
    float bspline_offset = 0;
    int image_dtype = 0;                           /* APRH_DTYPE_U16 / _F32 of the converted image */
    bool borrowed = false;                         /* owned by a batch plan     */
    aprh::Buf xz_end;                              /* u64[n_xz]                 */
    aprh::Buf y_vec;                               /* u16[total]                */
    /* optional stage outputs (context keep_intermediates) */
    aprh::Buf smoothed, grad_raw, lis_raw, grad, lis, levels;
    std::vector<aprh::Buf> pct;                    /* final particle-cell tree  */
    std::vector<int64_t> pct_dims;                 /* 3 per pct level           */
    /* tree (interior cells): GenInfo::init_tree levels 0..t_l_max + LinearAccess,
       built on first use */
    bool tree_built = false;
    int t_l_max = 0;
    std::vector<int64_t> t_y, t_x, t_z;
    std::vector<uint64_t> t_level_xz;
    uint64_t t_n_xz = 0, t_total = 0;
    aprh::Buf t_xz_end;                            /* u64[t_n_xz]               */
    aprh::Buf t_y_vec;                             /* u16[t_total]              */
    /* per-row index of 64-cell y blocks (APR and tree rows), built on first convolution */
    bool yidx_built = false;
    aprh::Buf yidx, t_yidx;                        /* u16                       */
    std::vector<int64_t> yidx_off, t_yidx_off;     /* per level                 */
    int top_tree_regular = -1;                     /* top tree level readable straight from the particle rows: -1 unknown */
};

namespace aprh {

/* level geometry of GenInfo::init (GenInfo.hpp:33-65) */
void geninfo_init(aprh_apr& a, int64_t y, int64_t x, int64_t z);

/* does ptr refer to device memory? */
bool is_device_ptr(const void* p);

/* copy helpers that accept host or device pointers on either side */
void copy_any(aprh_context* ctx, void* dst, const void* src, size_t bytes);

/* B-spline coefficients computed on the host (ComputeGradient.hpp:256-330) */
struct IIRCoeffs {
    float b1, b2, nf;
    int mL;
    float bc1[64], bc2[64], bc3[64], bc4[64];
};
IIRCoeffs iir_coeffs(float lambda, float tol, int64_t len);

/* ---- stage launchers (k_iir.hip) ---- */
void launch_minmax_offset(aprh_context* ctx, const uint16_t* img, size_t n, float lambda, float* d_offset, uint32_t* d_minmax);
void launch_offset_copy(aprh_context* ctx, const uint16_t* in, uint16_t* out, size_t n, const float* d_offset);
template <typename T>
void launch_iir_y(aprh_context* ctx, const T* in, T* out, int64_t y, int64_t x, int64_t z, const IIRCoeffs& c,
                  const float* d_offset, int* d_err);
/* u16 y pass that also finds the bspline offset: it runs with the offset most images get, gathers
   the input's (min, max) on the way, and runs again only when the true offset differs */
bool iir_y_fuses_offset(int64_t y, const void* in, const void* out);
/* pyr (optional): receives the first level of the input's mean pyramid (ceil(x/2) x ceil(z/2) rows
   of y/2, 16-byte aligned) */
void launch_iir_y_offset(aprh_context* ctx, const uint16_t* in, uint16_t* out, int64_t y, int64_t x, int64_t z, const IIRCoeffs& c,
                         float lambda, float* d_offset, uint32_t* d_mm, int* d_rerun, int* d_err, uint16_t* pyr);
template <typename T>
void launch_iir_outer(aprh_context* ctx, T* img, int64_t y, int64_t x, int64_t z, int axis, const IIRCoeffs& c, int* d_err);

/* ---- gradient / inverse B-spline / LIS (k_grad_lis.hip) ---- */
/* accum: take the existing grad values into the maximum (calc_bspline_fd_ds_mag semantics);
   the pipeline passes false on a buffer it would otherwise have zero-filled */
void launch_grad_ds_u16(aprh_context* ctx, const uint16_t* img, uint16_t* grad, float* mean, int64_t y, int64_t x, int64_t z,
                        float hx, float hy, float hz, bool accum);
void launch_inv_bspline(aprh_context* ctx, float* m, float* tmp, int64_t y, int64_t x, int64_t z, bool ay, bool ax, bool az);
/* the same passes alternating between a (input) and b; returns the buffer holding the result */
float* launch_inv_bspline_pp(aprh_context* ctx, float* a, float* b, int64_t y, int64_t x, int64_t z, bool ay, bool ax, bool az);
/* number of a -> b launches launch_inv_bspline_pp makes (its result is in b when odd) */
int inv_bspline_launches(bool ay, bool ax, bool az);
struct LisParams {
    int win[6];
    bool active[3];
    float var_rescale;
    int pad[3];
    int pitch_align;   /* row pitch of the padded array: yp rounded up to this many floats (0 or 1: yp) */
    bool fuse_levels = false;   /* launch_lis leaves out the last z mean; the levels launch runs it with its epilogue */
};
LisParams lis_params(const aprh_params& par, int64_t yd, int64_t xd, int64_t zd);
/* full local intensity scale over the reflect-padded array A (see lis_padded_elems) */
void launch_lis(aprh_context* ctx, const float* intensity, float* A, float* unused, int64_t yd, int64_t xd, int64_t zd,
                const LisParams& lp);
size_t lis_padded_elems(int64_t yd, int64_t xd, int64_t zd, const LisParams& lp);
/* whether the last z mean can run inside the levels launch (LisParams::fuse_levels) */
bool lis_can_fuse_levels(const LisParams& lp, int64_t yd, int64_t xd, int64_t zd);
void launch_unpad_rescale(aprh_context* ctx, const float* A, const LisParams& lp, int64_t yd, int64_t xd, int64_t zd, float* out);
/* constant-scale branch of get_local_intensity_scale, on the device */
void launch_constant_scale(aprh_context* ctx, float* m, size_t n);
void launch_u8_to_f32(aprh_context* ctx, const uint8_t* in, float* out, size_t n);
void launch_sat_mean(aprh_context* ctx, float* m, int64_t y, int64_t x, int64_t z, int axis, int off);
/* rescale + unpad of A, applyParameters and computeLevels (level map u8);
   optional outputs: sigma_raw (local_intensity_scale_step), sigma_out (after sigma_th), grad_keep */
template <typename G>
void launch_levels_padded(aprh_context* ctx, const float* A, const LisParams& lp, int64_t yd, int64_t xd, int64_t zd,
                          const float* inten, G* grad, float* sigma_raw, float* sigma_out, uint8_t* level,
                          const aprh_params& par, const float* d_offset, float mult, G* grad_keep);
/* gradient magnitude + 2x2x2 max + mean downsample of a float image (APRConverter<float>) */
void launch_grad_ds_f32(aprh_context* ctx, const float* img, float* grad, float* mean, int64_t y, int64_t x, int64_t z, float hx,
                        float hy, float hz, bool accum);

/* ---- pulling scheme / structure / sampling (k_structure.hip) ---- */
void build_structure(aprh_context* ctx, aprh_apr& a, const uint8_t* level_top, const aprh_params& par);
/* the same from a particle-cell tree already seeded (seeded != nullptr; pct_layout), or from
   level_top as above */
void build_structure_seeded(aprh_context* ctx, aprh_apr& a, const uint8_t* level_top, uint8_t* seeded, const aprh_params& par);
std::vector<int64_t> pct_layout(const aprh_apr& a);
/* fill_patch of every level for one z block (patch of zp planes starting at z0 - gl) */
void seed_patch(aprh_context* ctx, const aprh_apr& a, uint8_t* P, const std::vector<int64_t>& off, const uint8_t* level_top_patch,
                int64_t zp, int64_t z0, int64_t gl, int64_t gr);
/* ParticleData::sample_parts_from_img_blocked on a device image */
void sample_blocked_u16(aprh_context* ctx, const aprh_apr& a, const uint16_t* img, int64_t z_block_size, int64_t ghost_z,
                        uint16_t* parts);
void launch_sample_u16(aprh_context* ctx, const aprh_apr& a, const uint16_t* img, uint16_t* parts, uint16_t* pyr_ext = nullptr);
/* bytes of the mean pyramid (levels l_min .. l_max-1) sample_image builds, elements of `elem` bytes */
size_t sample_pyramid_bytes(const aprh_apr& a, size_t elem);
/* resolves the device address of the no-alloc scratch array on `device` (current) */
void noalloc_init(int device);
/* element offsets of the mean pyramid's levels l_min .. l_max-1 in one buffer; *total elements */
std::vector<int64_t> pyramid_layout(const aprh_apr& a, int64_t* total);
/* ParticleData<float>::sample_image of a float image (same pyramid, float means) */
void launch_sample_f32(aprh_context* ctx, const aprh_apr& a, const float* img, float* parts);

/* exclusive scan of n u32 counts into u64 offsets; returns the total (synchronises the stream) */
uint64_t exclusive_scan_u32(aprh_context* ctx, const uint32_t* in, uint64_t* out, int64_t n);
/* the same, total left in device memory at d_total, no synchronisation */
void exclusive_scan_u32_async(aprh_context* ctx, const uint32_t* in, uint64_t* out, int64_t n, uint64_t* d_total);
/* zero device memory with a kernel (graph-capture safe, unlike a memset node) */
void launch_zero(aprh_context* ctx, void* p, size_t bytes);
/* device-to-device copy with a kernel (recordable by a batch plan) */
void launch_copy(aprh_context* ctx, void* dst, const void* src, size_t bytes);

/* ---- particle operations (k_particles.hip) ---- */
struct HostStencil {
    int sy = 1, sx = 1, sz = 1;
    std::vector<float> w;   /* (z, x, y) order, y fastest */
};
/* APRStencil::get_downsampled_stencils (APRStencil.hpp:81-97) on the host */
std::vector<HostStencil> downsampled_stencils(const HostStencil& in, int nlevels, bool normalize);
/* throws APRH_ERR_INVALID_ARGUMENT unless xz_end / y_vec describe a.org's level geometry */
void validate_structure(aprh_context* ctx, const aprh_apr& a);
void ensure_tree(aprh_context* ctx, aprh_apr& a);
template <typename T>
void tree_mean(aprh_context* ctx, aprh_apr& a, const T* parts, float* tree);
/* tree_in: the tree values to use (APRFilter.hpp:1094-1100's tree_data); nullptr: the tree
   means of `in` (fill_tree_mean, APRFilter.hpp:1088) */
template <typename T>
void convolve(aprh_context* ctx, aprh_apr& a, const std::vector<HostStencil>& stencils, const T* in, float* out, bool reflect,
              const float* tree_in = nullptr);
/* 1-D factors (fy, fx, fz) of a rank-1 stencil (to 1e-6 of its largest weight), if it is one */
bool rank1_factors(const HostStencil& s, std::vector<float>& fy, std::vector<float>& fx, std::vector<float>& fz);
/* ---- level-tile convolution (k_conv.hip) ---- */
void ensure_yidx(aprh_context* ctx, aprh_apr& a);
template <typename T>
bool tree_mean_top(aprh_context* ctx, aprh_apr& a, const T* parts, float* tree);
template <typename T>
void conv_levels(aprh_context* ctx, aprh_apr& a, const std::vector<HostStencil>& stencils, bool sep_mode, const T* in, const float* tree,
                 float* out, bool reflect, int epi, const void* epi_in, bool epi_u16, Buf* part, int64_t* nparts);
void reduce_norm_parts(aprh_context* ctx, const double* part, int64_t nparts, double* d_out);
/* elementwise steps of APRNumerics::local_std: mode 0 p = float(u16 in), 1 p = p * p,
   2 p = sqrtf(max(p - m * m, 0)) */
void local_std_step(aprh_context* ctx, int mode, float* p, const void* in_or_m, uint64_t n);
/* elementwise steps of APRNumerics::richardson_lucy_tv: mode 0 normalise (p = g_y, q = g_x,
   r = g_z; q / r may be nullptr), 1 p = q + p, 2 p = p * q / (1 - reg * r) */
void rl_tv_step(aprh_context* ctx, int mode, float* p, float* q, float* r, float reg, uint64_t n);
void fill_f32(aprh_context* ctx, float* p, size_t n, float v);
template <typename T>
int richardson_lucy(aprh_context* ctx, aprh_apr& a, const T* in, float* u, const std::vector<HostStencil>& psf,
                    const std::vector<HostStencil>& psf_flipped, int max_iter, bool resume, double tol, double* last_rel);
/* piecewise-constant image of level `level` (l_max: APRReconstruction::reconstruct_constant);
   dtypes 0 = u16, 1 = f32 for particles (pdt) and image (idt) */
void reconstruct_any(aprh_context* ctx, aprh_apr& a, const void* parts, int pdt, int level, void* img, int idt);
/* level image at `level` (<= l_max) with pixels of finer particles set to `level` */
void reconstruct_level(aprh_context* ctx, aprh_apr& a, int level, void* img, int idt);
/* APRReconstruction::reconstruct_smooth of the whole image (scale3: y, x, z) */
void reconstruct_smooth(aprh_context* ctx, aprh_apr& a, const void* parts, int pdt, const float* scale3, void* img, int idt);
/* APRNumerics::face_neighbour_filter along dim (filter3: -face, centre, +face weights); in and
   out must differ */
void face_neighbour_filter(aprh_context* ctx, const aprh_apr& a, const void* in, int in_dtype, float* out, const float* filter3, int dim);
/* g = sqrtf(g*g [+ gx*gx] [+ gz*gz]) elementwise (gx / gz may be nullptr) */
void gradient_magnitude_combine(aprh_context* ctx, float* g, const float* gx, const float* gz, uint64_t n);
/* ReconPatch::check_limits: clamps the bounds; false when a range is empty */
bool patch_limits(const aprh_apr& a, aprh_patch& p);
/* patch of the constant (parts != nullptr) or level reconstruction; p already clamped */
void reconstruct_patch(aprh_context* ctx, aprh_apr& a, const void* parts, int pdt, const aprh_patch& p, void* img, int idt);

/* ---- synthetic data (k_synth.hip) ---- */
void synth_deconv_device(aprh_context* ctx, float* gt, uint16_t* noisy, int64_t y, int64_t x, int64_t z, uint64_t seed, int64_t n_blobs,
                         float smin, float smax, float amin, float amax, float background, const float* psf_sigma3, int taps);
void synth_device(aprh_context* ctx, uint16_t* out, int64_t y, int64_t x, int64_t z, uint64_t seed, int64_t n_blobs, float smin,
                  float smax, float amin, float amax, float background, int noise);

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

/* Division of 32-bit unsigned values by a divisor fixed at launch: a multiply-high, a
   subtract and two shifts (the round-up method of Granlund and Montgomery), exact for every
   32-bit dividend.  Grid-stride kernels use it to split a linear index into (y, x, z)
   instead of 64-bit division. */
struct FastDiv {
    uint32_t m = 0;
    int l = 0;
    FastDiv() = default;
    explicit FastDiv(uint64_t d) {
        if (d == 0 || d > 0xFFFFFFFFull) throw Error(APRH_ERR_UNSUPPORTED, "divisor outside 32 bits");
        while ((1ull << l) < d) ++l;
        m = (l == 0) ? 0u : (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
    }
    __device__ __forceinline__ uint32_t div(uint32_t n) const {
        if (l == 0) return n;
        const uint32_t t = __umulhi(m, n);
        return (t + ((n - t) >> 1)) >> (l - 1);
    }
};
/* a grid-stride loop over n items may use 32-bit indices */
inline bool fits32(int64_t n) { return n >= 0 && n < (int64_t)0x7FFFFFFF; }

/* record the start of stage s (the end of the previous one) when timing is on */
inline void mark(aprh_context* ctx, int s) {
    if (ctx->timing) {
        HIP_CHECK(hipEventRecord(ctx->ev[s], ctx->stream));
        ctx->ev_marked[s] = true;
    }
}
/* convert marked event pairs [first, last) into stage times */
void collect_times(aprh_context* ctx, int first, int last);
void reset_times(aprh_context* ctx, int first, int last);

void set_last_error(const std::string& s);

/* device copy of a caller buffer (no copy when it already is device memory) */
struct DevIn {
    const void* ptr = nullptr;
    Buf own;
    DevIn(aprh_context* ctx, const void* p, size_t bytes) {
        if (is_device_ptr(p)) {
            ptr = p;
        } else {
            own.reset(ctx->pool, bytes);
            HIP_CHECK(hipMemcpyAsync(own.ptr(), p, bytes, hipMemcpyHostToDevice, ctx->stream));
            ptr = own.ptr();
        }
    }
};

/* device output that is copied to the caller's host buffer on finish() */
struct DevOut {
    void* ptr = nullptr;
    void* host = nullptr;
    size_t bytes = 0;
    Buf own;
    DevOut(aprh_context* ctx, void* p, size_t b) : bytes(b) {
        if (is_device_ptr(p)) {
            ptr = p;
        } else {
            own.reset(ctx->pool, b);
            ptr = own.ptr();
            host = p;
        }
    }
    void finish(aprh_context* ctx) {
        if (host) HIP_CHECK(hipMemcpyAsync(host, ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
};

inline void check_ctx(aprh_context* ctx) {
    if (!ctx) throw Error(APRH_ERR_INVALID_ARGUMENT, "context is NULL");
    HIP_CHECK(hipSetDevice(ctx->device));
}


}  // namespace aprh

/* exception -> status translation used by every extern "C" entry point */
/* every entry point starts from a clean HIP error state, so an error left by an earlier,
   unrelated call is never reported by this one */
#define APRH_TRY                 \
    (void)hipGetLastError();     \
    try {
#define APRH_CATCH                                              \
    }                                                           \
    catch (aprh::Error & e) {                                   \
        aprh::set_last_error(e.what());                         \
        return e.code;                                          \
    }                                                           \
    catch (std::bad_alloc&) {                                   \
        aprh::set_last_error("host allocation failed");         \
        return APRH_ERR_OUT_OF_MEMORY;                          \
    }                                                           \
    catch (std::exception & e) {                                \
        aprh::set_last_error(e.what());                         \
        return APRH_ERR_HIP;                                    \
    }                                                           \
    return APRH_OK;

namespace aprh {

}  // namespace aprh

#endif
