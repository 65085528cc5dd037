/*
 * capi.hip -- extern "C" entry points of libaprh: the get_apr pipeline
 * (APRConverter<uint16_t>::get_apr, APRConverter.hpp:534-582), sampling, the
 * stage-level entry points and the APR accessors.
 */
#include <algorithm>
#include <cmath>

#include "aprh_internal.hpp"

using namespace aprh;

namespace {

void check_dims(int64_t y, int64_t x, int64_t z) {
    if (y <= 0 || x <= 0 || z <= 0) throw Error(APRH_ERR_BAD_DIMS, "dimensions must be positive");
    if (y > 65535 || x > 65535 || z > 65535) throw Error(APRH_ERR_UNSUPPORTED, "dimensions above 65535 (u16 y_vec / iterator limits)");
}

float level_mult(const aprh_params& par, int l_max) {
    /* computeLevels (LocalParticleCellSet.hpp:128-147) */
    const float min_dim = std::min(par.dy, std::min(par.dx, par.dz));
    const float level_factor = (float)(std::pow(2.0, (double)l_max) * (double)min_dim);
    return level_factor / par.rel_error;
}

void keep_copy(aprh_context* ctx, Buf& dst, const void* src, size_t bytes) {
    dst.reset(ctx->pool, bytes);
    HIP_CHECK(hipMemcpyAsync(dst.ptr(), src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
}

/* levels + pulling scheme + structure from a downsampled gradient, an unpadded
   scale (A with zero pads, var_rescale 1) or the padded LIS array */
template <typename G>
void solve(aprh_context* ctx, aprh_apr& a, const aprh_params& par, const float* A, const LisParams& lp, const float* inten,
           G* grad, const float* d_off) {
    const int64_t yd = cdiv(a.org[0], 2), xd = cdiv(a.org[1], 2), zd = cdiv(a.org[2], 2);
    const size_t nd = (size_t)(yd * xd * zd);
    Buf level(ctx->pool, nd);
    const bool keep = ctx->keep_intermediates != 0;
    float* sig_raw = nullptr;
    float* sig = nullptr;
    G* gk = nullptr;
    if (keep) {
        a.lis_raw.reset(ctx->pool, nd * 4);
        a.lis.reset(ctx->pool, nd * 4);
        a.grad.reset(ctx->pool, nd * sizeof(G));
        sig_raw = a.lis_raw.as<float>();
        sig = a.lis.as<float>();
        gk = a.grad.as<G>();
    }
    BlockReq* br = ctx->block_req;
    launch_levels_padded<G>(ctx, A, lp, yd, xd, zd, inten, grad, sig_raw, sig, level.as<uint8_t>(), par, d_off,
                            level_mult(par, br ? br->l_max : a.l_max), gk);
    if (br) {
        /* a z block of a blocked conversion: seed the global particle-cell tree (fill_patch) */
        if (br->l_max > 2) seed_patch(ctx, *br->global, br->pct, br->off, level.as<uint8_t>(), a.org[2], br->z0, br->gl, br->gr);
        return;
    }
    if (keep) {
        /* levels as float, like local_scale_temp after computeLevels */
        a.levels.reset(ctx->pool, nd * 4);
        launch_u8_to_f32(ctx, level.as<uint8_t>(), a.levels.as<float>(), nd);
    }
    build_structure(ctx, a, level.as<uint8_t>(), par);
}

template <typename G>
void lis_and_solve(aprh_context* ctx, aprh_apr* a, const aprh_params& par, Buf& mean, G* grad, float* d_off, const int* d_err);

void get_apr_u16(aprh_context* ctx, const uint16_t* image, int64_t y, int64_t x, int64_t z, const aprh_params& par, aprh_apr* a) {
    reset_times(ctx, ST_OFFSET, ST_SAMPLE);
    const size_t n = (size_t)(y * x * z);
    const int64_t yd = cdiv(y, 2), xd = cdiv(x, 2), zd = cdiv(z, 2);
    const size_t nd = (size_t)(yd * xd * zd);
    hipStream_t st = ctx->stream;
    DevIn in(ctx, image, n * 2);
    if (ctx->sample_req) ctx->sample_req->img = (const uint16_t*)in.ptr;
    geninfo_init(*a, y, x, z);
    mark(ctx, ST_OFFSET);
    /* offset, min/max and overflow words: the context's, or a recorded volume's own */
    Buf own_flags;
    if (ctx->recorder) own_flags.reset(ctx->pool, 256);
    int* flags = ctx->recorder ? own_flags.as<int>() : ctx->d_flags;
    float* d_off = (float*)flags;
    uint32_t* d_mm = (uint32_t*)flags + 2;
    int* d_err = flags + 4;
    int* d_rerun = flags + 8;
    launch_zero(ctx, d_err, 3 * sizeof(int));
    Buf img(ctx->pool, n * 2);
    uint16_t* I = img.as<uint16_t>();
    const float tol = 0.0001f;
    if (par.lambda > 0 && y > 2 && iir_y_fuses_offset(y, in.ptr, I)) {
        /* computeL's bspline offset (APRConverter.hpp:213-224) found by the y pass itself */
        mark(ctx, ST_SMOOTH_Y);
        /* with sampling requested, the pass also writes the first level of the image's mean pyramid */
        uint16_t* pyr_top = nullptr;
        SampleReq* sr = ctx->sample_req;
        if (sr && !ctx->recorder && a->l_max - 1 >= a->l_min && a->l_max > 2) {
            int64_t tot = 0;
            sr->pyr_off = pyramid_layout(*a, &tot);
            sr->pyr.reset(ctx->pool, (size_t)std::max<int64_t>(tot, 1) * 2);
            pyr_top = sr->pyr.as<uint16_t>() + sr->pyr_off[a->l_max - 1];
            sr->top_built = true;
        }
        launch_iir_y_offset(ctx, (const uint16_t*)in.ptr, I, y, x, z, iir_coeffs(par.lambda, tol, y), par.lambda, d_off, d_mm, d_rerun,
                            d_err + 0, pyr_top);
    } else {
        /* computeL: bspline offset (APRConverter.hpp:213-224) */
        launch_minmax_offset(ctx, (const uint16_t*)in.ptr, n, par.lambda, d_off, d_mm);
        mark(ctx, ST_SMOOTH_Y);
        /* smoothing (ComputeGradient.hpp:173-205) with the offset copy fused into the first pass */
        if (par.lambda > 0 && y > 2) {
            launch_iir_y<uint16_t>(ctx, (const uint16_t*)in.ptr, I, y, x, z, iir_coeffs(par.lambda, tol, y), d_off, d_err + 0);
        } else {
            launch_offset_copy(ctx, (const uint16_t*)in.ptr, I, n, d_off);
        }
    }
    mark(ctx, ST_SMOOTH_X);
    if (par.lambda > 0 && x > 2) launch_iir_outer<uint16_t>(ctx, I, y, x, z, 1, iir_coeffs(par.lambda, tol, x), d_err + 1);
    mark(ctx, ST_SMOOTH_Z);
    if (par.lambda > 0 && z > 2) launch_iir_outer<uint16_t>(ctx, I, y, x, z, 2, iir_coeffs(par.lambda, tol, z), d_err + 2);
    mark(ctx, ST_GRADIENT);
    /* gradient magnitude + 2x2x2 max (ComputeGradient.hpp:92-93), mean downsample and inverse B-spline (:95-118) */
    Buf grad(ctx->pool, nd * 2);
    /* the inverse B-spline runs tmp -> mean: the gradient kernel writes the downsampled mean
       where an odd number of launches ends in mean (no copy back) */
    const int npass = par.lambda > 0 ? inv_bspline_launches(y > 2, x > 2, z > 2) : 0;
    Buf mean(ctx->pool, nd * 4), tmp;
    if (npass) tmp.reset(ctx->pool, nd * 4);
    float* mean_first = (npass & 1) ? tmp.as<float>() : mean.as<float>();
    launch_grad_ds_u16(ctx, I, grad.as<uint16_t>(), mean_first, y, x, z, par.dx, par.dy, par.dz, false);
    /* the overflow counters are read with the B-spline offset at the end of the call (one host
       synchronisation instead of a stall between the gradient and the inverse B-spline) */
    if (ctx->async_mode) launch_copy(ctx, ctx->slot_err, d_err, 3 * sizeof(int));
    if (ctx->keep_intermediates) {
        keep_copy(ctx, a->smoothed, I, n * 2);
        keep_copy(ctx, a->grad_raw, grad.ptr(), nd * 2);
    }
    img.release();
    mark(ctx, ST_INV_BSPLINE);
    if (npass) launch_inv_bspline_pp(ctx, mean_first, (npass & 1) ? mean.as<float>() : tmp.as<float>(), yd, xd, zd, y > 2, x > 2, z > 2);
    tmp.release();
    lis_and_solve<uint16_t>(ctx, a, par, mean, grad.as<uint16_t>(), d_off, d_err);
}

/* local intensity scale, applyParameters, solveForAPR, generateDatastructures */
template <typename G>
void lis_and_solve(aprh_context* ctx, aprh_apr* a, const aprh_params& par, Buf& mean, G* grad, float* d_off, const int* d_err) {
    hipStream_t st = ctx->stream;
    const int64_t yd = cdiv(a->org[0], 2), xd = cdiv(a->org[1], 2), zd = cdiv(a->org[2], 2);
    const size_t nd = (size_t)(yd * xd * zd);
    mark(ctx, ST_LIS);
    /* local intensity scale (LocalIntensityScale.hpp:26-150) */
    LisParams lp = lis_params(par, yd, xd, zd);
    const int nact = (int)lp.active[0] + (int)lp.active[1] + (int)lp.active[2];
    Buf A;
    const float* Ap;
    LisParams lpa = lp;
    if (par.constant_intensity_scale || nact == 0) {
        A.reset(ctx->pool, nd * 4);
        launch_copy(ctx, A.ptr(), mean.ptr(), nd * 4);
        launch_constant_scale(ctx, A.as<float>(), nd);
        lpa.pad[0] = lpa.pad[1] = lpa.pad[2] = 0;
        lpa.pitch_align = 1;
        lpa.var_rescale = 1.0f;
        Ap = A.as<float>();
    } else {
        A.reset(ctx->pool, lis_padded_elems(yd, xd, zd, lp) * 4);
        /* the levels launch runs the last z mean (its epilogue also keeps the intermediates) */
        lp.fuse_levels = lis_can_fuse_levels(lp, yd, xd, zd);
        lpa = lp;
        launch_lis(ctx, mean.as<float>(), A.as<float>(), nullptr, yd, xd, zd, lp);
        Ap = A.as<float>();
    }
    /* applyParameters + solveForAPR + generateDatastructures */
    mark(ctx, ST_LEVELS);
    solve<G>(ctx, *a, par, Ap, lpa, mean.as<float>(), grad, d_off);
    mark(ctx, ST_SAMPLE);
    if (ctx->async_mode) {
        launch_copy(ctx, ctx->slot_offset, d_off, 4);
        return;
    }
    HIP_CHECK(hipMemcpyAsync(&a->bspline_offset, d_off, 4, hipMemcpyDeviceToHost, st));
    int herr[3] = {0, 0, 0};
    if (d_err) HIP_CHECK(hipMemcpyAsync(herr, d_err, sizeof(herr), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    collect_times(ctx, ST_OFFSET, ST_SAMPLE);
    /* ComputeGradient.hpp:409-412: the u16 smoothing's range check */
    static const char* axis_name[3] = {"y", "x", "z"};
    for (int i = 0; i < 3; ++i)
        if (herr[i])
            throw Error(APRH_ERR_OVERFLOW, std::string("integer under-/overflow encountered in ComputeGradient::bspline_filt_rec_") +
                                               axis_name[i] +
                                               " - try squashing the input image to a narrower range or use APRConverter<float>");
}

/* APRConverter<float>::get_apr: no B-spline offset (APRConverter.hpp:218-220), float
   smoothing without rounding, float gradient (grad_temp is PixelData<float>) */
void get_apr_f32(aprh_context* ctx, const float* image, int64_t y, int64_t x, int64_t z, const aprh_params& par, aprh_apr* a) {
    reset_times(ctx, ST_OFFSET, ST_SAMPLE);
    const size_t n = (size_t)(y * x * z);
    const int64_t yd = cdiv(y, 2), xd = cdiv(x, 2), zd = cdiv(z, 2);
    const size_t nd = (size_t)(yd * xd * zd);
    hipStream_t st = ctx->stream;
    DevIn in(ctx, image, n * 4);
    geninfo_init(*a, y, x, z);
    a->image_dtype = APRH_DTYPE_F32;
    a->bspline_offset = 0;
    mark(ctx, ST_OFFSET);
    float* d_off = (float*)ctx->d_flags;
    int* d_err = ctx->d_flags + 4;
    launch_zero(ctx, ctx->d_flags, 8 * sizeof(int));   /* offset 0.0f, error counters 0 */
    mark(ctx, ST_SMOOTH_Y);
    Buf img(ctx->pool, n * 4);
    float* I = img.as<float>();
    const float tol = 0.0001f;
    if (par.lambda > 0 && y > 2)
        launch_iir_y<float>(ctx, (const float*)in.ptr, I, y, x, z, iir_coeffs(par.lambda, tol, y), nullptr, d_err + 0);
    else
        HIP_CHECK(hipMemcpyAsync(I, in.ptr, n * 4, hipMemcpyDeviceToDevice, st));
    mark(ctx, ST_SMOOTH_X);
    if (par.lambda > 0 && x > 2) launch_iir_outer<float>(ctx, I, y, x, z, 1, iir_coeffs(par.lambda, tol, x), d_err + 1);
    mark(ctx, ST_SMOOTH_Z);
    if (par.lambda > 0 && z > 2) launch_iir_outer<float>(ctx, I, y, x, z, 2, iir_coeffs(par.lambda, tol, z), d_err + 2);
    mark(ctx, ST_GRADIENT);
    Buf grad(ctx->pool, nd * 4);
    const int npass = par.lambda > 0 ? inv_bspline_launches(y > 2, x > 2, z > 2) : 0;
    Buf mean(ctx->pool, nd * 4), tmp;
    if (npass) tmp.reset(ctx->pool, nd * 4);
    float* mean_first = (npass & 1) ? tmp.as<float>() : mean.as<float>();
    launch_grad_ds_f32(ctx, I, grad.as<float>(), mean_first, y, x, z, par.dx, par.dy, par.dz, false);
    if (ctx->keep_intermediates) {
        keep_copy(ctx, a->smoothed, I, n * 4);
        keep_copy(ctx, a->grad_raw, grad.ptr(), nd * 4);
    }
    img.release();
    mark(ctx, ST_INV_BSPLINE);
    if (npass) launch_inv_bspline_pp(ctx, mean_first, (npass & 1) ? mean.as<float>() : tmp.as<float>(), yd, xd, zd, y > 2, x > 2, z > 2);
    tmp.release();
    lis_and_solve<float>(ctx, a, par, mean, grad.as<float>(), d_off, nullptr);
}

}  // namespace

namespace aprh {
/* the u16 pipeline for callers in other translation units (the batched plan) */
void get_apr_u16_dispatch(aprh_context* ctx, const uint16_t* image, int64_t y, int64_t x, int64_t z, const aprh_params& par,
                          aprh_apr* a) {
    get_apr_u16(ctx, image, y, x, z, par, a);
}
}  // namespace aprh

extern "C" {

int aprh_get_apr_u16(aprh_context* ctx, const uint16_t* image, int64_t y, int64_t x, int64_t z, const aprh_params* par,
                     aprh_apr** out) {
    APRH_TRY
    check_ctx(ctx);
    check_dims(y, x, z);
    if (!image || !par || !out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    aprh_apr* a = new aprh_apr();
    a->ctx = ctx;
    try {
        get_apr_u16(ctx, image, y, x, z, *par, a);
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        delete a;
        throw;
    }
    *out = a;
    APRH_CATCH
}

int aprh_get_apr_sample_u16(aprh_context* ctx, const uint16_t* image, int64_t y, int64_t x, int64_t z, const aprh_params* par,
                            aprh_apr** out, uint16_t* parts, uint64_t parts_capacity, int32_t* sampled) {
    APRH_TRY
    check_ctx(ctx);
    check_dims(y, x, z);
    if (!image || !par || !out || !sampled) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (parts_capacity > 0 && !parts) throw Error(APRH_ERR_INVALID_ARGUMENT, "parts is NULL with a nonzero capacity");
    *sampled = 0;
    if (parts && !is_device_ptr(parts)) throw Error(APRH_ERR_INVALID_ARGUMENT, "parts must be device memory");
    aprh_apr* a = new aprh_apr();
    a->ctx = ctx;
    SampleReq req;
    req.parts = parts;
    req.capacity = parts ? parts_capacity : 0;
    try {
        ctx->sample_req = &req;
        get_apr_u16(ctx, image, y, x, z, *par, a);
        ctx->sample_req = nullptr;
        if (!req.done && parts && a->total <= parts_capacity) {
            /* conversions whose structure is not built by the row fill (l_max <= 2) */
            DevIn in(ctx, image, (size_t)(y * x * z) * 2);
            launch_sample_u16(ctx, *a, (const uint16_t*)in.ptr, parts);
            req.done = true;
        }
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        ctx->sample_req = nullptr;
        (void)hipStreamSynchronize(ctx->stream);
        delete a;
        throw;
    }
    *sampled = req.done ? 1 : 0;
    *out = a;
    APRH_CATCH
}

/* APRConverterBatch<uint16_t>::get_apr (APRConverterBatch.hpp:111-175) on an image in host or
   device memory: z blocks with ghost slabs, each through the pixel pipeline on its own, seeds
   merged into one particle-cell tree, one pulling scheme and linear structure */
int aprh_get_apr_by_block_u16(aprh_context* ctx, const uint16_t* image, int64_t y, int64_t x, int64_t z, const aprh_params* par,
                              int64_t z_block_size, int64_t ghost_z, aprh_apr** out) {
    APRH_TRY
    check_ctx(ctx);
    check_dims(y, x, z);
    if (!image || !par || !out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (z_block_size <= 0 || ghost_z < 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "z_block_size must be positive and ghost_z >= 0");
    aprh_apr* a = new aprh_apr();
    a->ctx = ctx;
    const int keep = ctx->keep_intermediates;
    BlockReq req;
    try {
        geninfo_init(*a, y, x, z);
        req.l_max = a->l_max;
        req.global = a;
        req.off = a->l_max > 2 ? pct_layout(*a) : std::vector<int64_t>();
        Buf pct;
        if (a->l_max > 2) {
            pct.reset(ctx->pool, (size_t)std::max<int64_t>(req.off[a->l_max], 8));
            launch_zero(ctx, pct.ptr(), (size_t)req.off[a->l_max]);
            req.pct = pct.as<uint8_t>();
        }
        ctx->keep_intermediates = 0;
        const int64_t plane = y * x;
        const int64_t nb = std::max<int64_t>(z / z_block_size, 1);
        float offset = 0;
        for (int64_t b = 0; b < nb; ++b) {
            req.z0 = b * z_block_size;
            const int64_t zf = (b == nb - 1) ? z : (b + 1) * z_block_size;
            req.gl = std::min(req.z0, ghost_z);
            req.gr = std::min(z - zf, ghost_z);
            const int64_t zb = req.z0 - req.gl, zp = zf + req.gr - zb;
            aprh_apr patch;
            patch.ctx = ctx;
            ctx->block_req = &req;
            get_apr_u16(ctx, image + zb * plane, y, x, zp, *par, &patch);
            ctx->block_req = nullptr;
            offset = patch.bspline_offset;
        }
        ctx->keep_intermediates = keep;
        build_structure_seeded(ctx, *a, nullptr, req.pct, *par);
        a->bspline_offset = offset;
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        ctx->block_req = nullptr;
        ctx->keep_intermediates = keep;
        (void)hipStreamSynchronize(ctx->stream);
        delete a;
        throw;
    }
    *out = a;
    APRH_CATCH
}

/* ParticleData<uint16_t>::sample_parts_from_img_blocked (ParticleData.hpp:140-141, 292-362) on an
   image in host or device memory */
int aprh_sample_image_blocked_u16(aprh_context* ctx, const aprh_apr* a, const uint16_t* image, int64_t z_block_size, int64_t ghost_z,
                                  uint16_t* parts) {
    APRH_TRY
    check_ctx(ctx);
    if (!a || !image || !parts) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (z_block_size <= 0 || ghost_z < 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "z_block_size must be positive and ghost_z >= 0");
    const size_t n = (size_t)(a->org[0] * a->org[1] * a->org[2]);
    DevIn in(ctx, image, n * 2);
    DevOut o(ctx, parts, std::max<uint64_t>(a->total, 1) * 2);
    sample_blocked_u16(ctx, *a, (const uint16_t*)in.ptr, z_block_size, ghost_z, (uint16_t*)o.ptr);
    o.finish(ctx);
    APRH_CATCH
}

int aprh_get_apr_f32(aprh_context* ctx, const float* image, int64_t y, int64_t x, int64_t z, const aprh_params* par, aprh_apr** out) {
    APRH_TRY
    check_ctx(ctx);
    check_dims(y, x, z);
    if (!image || !par || !out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    aprh_apr* a = new aprh_apr();
    a->ctx = ctx;
    try {
        get_apr_f32(ctx, image, y, x, z, *par, a);
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        delete a;
        throw;
    }
    *out = a;
    APRH_CATCH
}

int aprh_apr_from_grad_lis(aprh_context* ctx, const uint16_t* grad_ds, const float* lis_ds, const float* intensity_ds, int64_t y,
                           int64_t x, int64_t z, const aprh_params* par, float bspline_offset, aprh_apr** out) {
    APRH_TRY
    check_ctx(ctx);
    if (!grad_ds || !lis_ds || !intensity_ds || !par || !out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    check_dims(y, x, z);
    const size_t nd = (size_t)(cdiv(y, 2) * cdiv(x, 2) * cdiv(z, 2));
    aprh_apr* a = new aprh_apr();
    a->ctx = ctx;
    try {
        geninfo_init(*a, y, x, z);
        a->bspline_offset = bspline_offset;
        DevIn g(ctx, grad_ds, nd * 2), l(ctx, lis_ds, nd * 4), it(ctx, intensity_ds, nd * 4);
        Buf gcopy(ctx->pool, nd * 2);
        HIP_CHECK(hipMemcpyAsync(gcopy.ptr(), g.ptr, nd * 2, hipMemcpyDeviceToDevice, ctx->stream));
        float* d_off = (float*)ctx->d_flags;
        HIP_CHECK(hipMemcpyAsync(d_off, &bspline_offset, 4, hipMemcpyHostToDevice, ctx->stream));
        LisParams lp{};
        lp.var_rescale = 1.0f;
        solve(ctx, *a, *par, (const float*)l.ptr, lp, (const float*)it.ptr, gcopy.as<uint16_t>(), d_off);
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        delete a;
        throw;
    }
    *out = a;
    APRH_CATCH
}

int aprh_apr_info_get(const aprh_apr* a, aprh_apr_info* info) {
    APRH_TRY
    if (!a || !info) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    info->l_min = a->l_min;
    info->l_max = a->l_max;
    for (int i = 0; i < 3; ++i) info->org_dims[i] = a->org[i];
    info->total_number_particles = a->total;
    info->xz_end_size = a->n_xz;
    info->level_xz_size = a->level_xz_vec.size();
    info->bspline_offset = a->bspline_offset;
    info->image_dtype = a->image_dtype;
    APRH_CATCH
}

int aprh_apr_copy_structure(const aprh_apr* a, uint64_t* level_xz_vec, uint64_t* xz_end_vec, uint16_t* y_vec) {
    APRH_TRY
    if (!a) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL apr");
    check_ctx(a->ctx);
    if (level_xz_vec) std::copy(a->level_xz_vec.begin(), a->level_xz_vec.end(), level_xz_vec);
    if (xz_end_vec) copy_any(a->ctx, xz_end_vec, a->xz_end.ptr(), a->n_xz * 8);
    if (y_vec) copy_any(a->ctx, y_vec, a->y_vec.ptr(), a->total * 2);
    HIP_CHECK(hipStreamSynchronize(a->ctx->stream));
    APRH_CATCH
}

int aprh_apr_device_structure(const aprh_apr* a, const uint64_t** xz_end_vec, const uint16_t** y_vec) {
    APRH_TRY
    if (!a) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL apr");
    if (xz_end_vec) *xz_end_vec = a->xz_end.as<uint64_t>();
    if (y_vec) *y_vec = a->y_vec.as<uint16_t>();
    APRH_CATCH
}

int aprh_apr_intermediate(const aprh_apr* a, int which, void* dst) {
    APRH_TRY
    if (!a || !dst) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    check_ctx(a->ctx);
    const aprh::Buf* b = nullptr;
    switch (which) {
        case 0: b = &a->smoothed; break;
        case 1: b = &a->grad_raw; break;
        case 2: b = &a->lis_raw; break;
        case 3: b = &a->grad; break;
        case 4: b = &a->lis; break;
        case 5: b = &a->levels; break;
        default: throw Error(APRH_ERR_INVALID_ARGUMENT, "unknown intermediate");
    }
    if (!b->ptr()) throw Error(APRH_ERR_INVALID_ARGUMENT, "intermediate not kept (aprh_context_set_keep_intermediates)");
    copy_any(a->ctx, dst, b->ptr(), b->bytes());
    HIP_CHECK(hipStreamSynchronize(a->ctx->stream));
    APRH_CATCH
}

int aprh_apr_particle_cell_tree(const aprh_apr* a, int level, uint8_t* dst, int64_t* dims3) {
    APRH_TRY
    if (!a) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL apr");
    check_ctx(a->ctx);
    if (level < 0 || level >= (int)a->pct.size() || !a->pct[level].ptr())
        throw Error(APRH_ERR_INVALID_ARGUMENT, "particle-cell tree level not kept");
    if (dims3)
        for (int i = 0; i < 3; ++i) dims3[i] = a->pct_dims[3 * level + i];
    if (dst) {
        copy_any(a->ctx, dst, a->pct[level].ptr(),
                 (size_t)(a->pct_dims[3 * level] * a->pct_dims[3 * level + 1] * a->pct_dims[3 * level + 2]));
        HIP_CHECK(hipStreamSynchronize(a->ctx->stream));
    }
    APRH_CATCH
}

int aprh_apr_destroy(aprh_apr* a) {
    APRH_TRY
    if (!a) return APRH_OK;
    if (a->borrowed) throw Error(APRH_ERR_INVALID_ARGUMENT, "this APR belongs to a batch plan (aprh_batch_destroy frees it)");
    if (a->ctx) {
        (void)hipSetDevice(a->ctx->device);
        (void)hipStreamSynchronize(a->ctx->stream);
    }
    delete a;
    APRH_CATCH
}

int aprh_sample_image_u16(aprh_context* ctx, const aprh_apr* a, const uint16_t* image, uint16_t* parts) {
    APRH_TRY
    check_ctx(ctx);
    if (!a || !image || !parts) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t n = (size_t)(a->org[0] * a->org[1] * a->org[2]);
    reset_times(ctx, ST_SAMPLE, ST_COUNT);
    DevIn in(ctx, image, n * 2);
    DevOut out(ctx, parts, std::max<uint64_t>(a->total, 1) * 2);
    mark(ctx, ST_SAMPLE);
    launch_sample_u16(ctx, *a, (const uint16_t*)in.ptr, (uint16_t*)out.ptr);
    mark(ctx, ST_COUNT);
    out.finish(ctx);
    collect_times(ctx, ST_SAMPLE, ST_COUNT);
    APRH_CATCH
}

int aprh_sample_image_f32(aprh_context* ctx, const aprh_apr* a, const float* image, float* parts) {
    APRH_TRY
    check_ctx(ctx);
    if (!a || !image || !parts) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t n = (size_t)(a->org[0] * a->org[1] * a->org[2]);
    DevIn in(ctx, image, n * 4);
    DevOut out(ctx, parts, std::max<uint64_t>(a->total, 1) * 4);
    launch_sample_f32(ctx, *a, (const float*)in.ptr, (float*)out.ptr);
    out.finish(ctx);
    APRH_CATCH
}

/* ---------------- stage-level entry points ---------------- */

int aprh_bspline_filt_rec_u16(aprh_context* ctx, uint16_t* image, int64_t y, int64_t x, int64_t z, int axis, float lambda,
                              float tol) {
    APRH_TRY
    check_ctx(ctx);
    if (!image) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL image");
    check_dims(y, x, z);
    const int64_t len = axis == 0 ? y : axis == 1 ? x : z;
    if (len < 2) throw Error(APRH_ERR_BAD_DIMS, "filtered axis must have at least 2 samples");
    const size_t n = (size_t)(y * x * z);
    DevOut io(ctx, image, n * 2);
    if (io.host) HIP_CHECK(hipMemcpyAsync(io.ptr, image, n * 2, hipMemcpyHostToDevice, ctx->stream));
    int* d_err = ctx->d_flags + 4;
    HIP_CHECK(hipMemsetAsync(d_err, 0, 4, ctx->stream));
    IIRCoeffs c = iir_coeffs(lambda, tol, len);
    if (axis == 0) launch_iir_y<uint16_t>(ctx, (const uint16_t*)io.ptr, (uint16_t*)io.ptr, y, x, z, c, nullptr, d_err);
    else launch_iir_outer<uint16_t>(ctx, (uint16_t*)io.ptr, y, x, z, axis, c, d_err);
    int herr = 0;
    HIP_CHECK(hipMemcpyAsync(&herr, d_err, 4, hipMemcpyDeviceToHost, ctx->stream));
    io.finish(ctx);
    if (herr) throw Error(APRH_ERR_OVERFLOW, "integer under-/overflow encountered in ComputeGradient::bspline_filt_rec");
    APRH_CATCH
}

int aprh_bspline_filt_rec_f32(aprh_context* ctx, float* image, int64_t y, int64_t x, int64_t z, int axis, float lambda, float tol) {
    APRH_TRY
    check_ctx(ctx);
    if (!image) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL image");
    check_dims(y, x, z);
    const int64_t len = axis == 0 ? y : axis == 1 ? x : z;
    if (len < 2) throw Error(APRH_ERR_BAD_DIMS, "filtered axis must have at least 2 samples");
    const size_t n = (size_t)(y * x * z);
    DevOut io(ctx, image, n * 4);
    if (io.host) HIP_CHECK(hipMemcpyAsync(io.ptr, image, n * 4, hipMemcpyHostToDevice, ctx->stream));
    int* d_err = ctx->d_flags + 4;
    IIRCoeffs c = iir_coeffs(lambda, tol, len);
    if (axis == 0) launch_iir_y<float>(ctx, (const float*)io.ptr, (float*)io.ptr, y, x, z, c, nullptr, d_err);
    else launch_iir_outer<float>(ctx, (float*)io.ptr, y, x, z, axis, c, d_err);
    io.finish(ctx);
    APRH_CATCH
}

int aprh_inv_bspline_f32(aprh_context* ctx, float* image, int64_t y, int64_t x, int64_t z, int axis) {
    APRH_TRY
    check_ctx(ctx);
    if (!image) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL image");
    check_dims(y, x, z);
    const int64_t len = axis == 0 ? y : axis == 1 ? x : z;
    if (len < 2) throw Error(APRH_ERR_BAD_DIMS, "filtered axis must have at least 2 samples");
    const size_t n = (size_t)(y * x * z);
    DevOut io(ctx, image, n * 4);
    if (io.host) HIP_CHECK(hipMemcpyAsync(io.ptr, image, n * 4, hipMemcpyHostToDevice, ctx->stream));
    Buf tmp(ctx->pool, n * 4);
    launch_inv_bspline(ctx, (float*)io.ptr, tmp.as<float>(), y, x, z, axis == 0, axis == 1, axis == 2);
    io.finish(ctx);
    APRH_CATCH
}

int aprh_gradient_ds_u16(aprh_context* ctx, const uint16_t* image, uint16_t* grad_ds, float* mean_ds, int64_t y, int64_t x, int64_t z,
                         float hx, float hy, float hz) {
    APRH_TRY
    check_ctx(ctx);
    if (!image || !grad_ds) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    check_dims(y, x, z);
    const size_t n = (size_t)(y * x * z);
    const size_t nd = (size_t)(cdiv(y, 2) * cdiv(x, 2) * cdiv(z, 2));
    DevIn in(ctx, image, n * 2);
    DevOut g(ctx, grad_ds, nd * 2);
    if (g.host) HIP_CHECK(hipMemcpyAsync(g.ptr, grad_ds, nd * 2, hipMemcpyHostToDevice, ctx->stream));
    if (mean_ds) {
        DevOut m(ctx, mean_ds, nd * 4);
        launch_grad_ds_u16(ctx, (const uint16_t*)in.ptr, (uint16_t*)g.ptr, (float*)m.ptr, y, x, z, hx, hy, hz, true);
        m.finish(ctx);
    } else {
        launch_grad_ds_u16(ctx, (const uint16_t*)in.ptr, (uint16_t*)g.ptr, nullptr, y, x, z, hx, hy, hz, true);
    }
    g.finish(ctx);
    APRH_CATCH
}

int aprh_sat_mean_f32(aprh_context* ctx, float* image, int64_t y, int64_t x, int64_t z, int axis, int offset) {
    APRH_TRY
    check_ctx(ctx);
    if (!image) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL image");
    check_dims(y, x, z);
    if (offset < 1 || offset > 6) throw Error(APRH_ERR_UNSUPPORTED, "offset must be 1..6");
    const size_t n = (size_t)(y * x * z);
    DevOut io(ctx, image, n * 4);
    if (io.host) HIP_CHECK(hipMemcpyAsync(io.ptr, image, n * 4, hipMemcpyHostToDevice, ctx->stream));
    launch_sat_mean(ctx, (float*)io.ptr, y, x, z, axis, offset);
    io.finish(ctx);
    APRH_CATCH
}

int aprh_local_intensity_scale(aprh_context* ctx, const float* intensity_ds, float* scale_ds, int64_t y, int64_t x, int64_t z,
                               const aprh_params* par) {
    APRH_TRY
    check_ctx(ctx);
    if (!intensity_ds || !scale_ds || !par) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    check_dims(y, x, z);
    const size_t n = (size_t)(y * x * z);
    DevIn in(ctx, intensity_ds, n * 4);
    DevOut out(ctx, scale_ds, n * 4);
    LisParams lp = lis_params(*par, y, x, z);
    const int nact = (int)lp.active[0] + (int)lp.active[1] + (int)lp.active[2];
    if (par->constant_intensity_scale || nact == 0) {
        HIP_CHECK(hipMemcpyAsync(out.ptr, in.ptr, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
        launch_constant_scale(ctx, (float*)out.ptr, n);
    } else {
        Buf A(ctx->pool, lis_padded_elems(y, x, z, lp) * 4);
        launch_lis(ctx, (const float*)in.ptr, A.as<float>(), nullptr, y, x, z, lp);
        launch_unpad_rescale(ctx, A.as<float>(), lp, y, x, z, (float*)out.ptr);
    }
    out.finish(ctx);
    APRH_CATCH
}

}  // extern "C"
