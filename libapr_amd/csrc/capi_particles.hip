/*
 * capi_particles.hip -- extern "C" entry points for operations on the
 * particles of an APR: structure import, tree, fill_tree_mean, stencils,
 * convolution, Richardson-Lucy and piecewise-constant reconstruction.
 */
#include <algorithm>
#include <cmath>

#include "aprh_internal.hpp"

using namespace aprh;

namespace {

size_t dtype_size(int dt) {
    if (dt == APRH_DTYPE_U16) return 2;
    if (dt == APRH_DTYPE_F32) return 4;
    throw Error(APRH_ERR_INVALID_ARGUMENT, "dtype must be APRH_DTYPE_U16 or APRH_DTYPE_F32");
}

HostStencil host_stencil(const float* w, int sy, int sx, int sz) {
    if (!w) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL stencil");
    if (sy < 1 || sx < 1 || sz < 1) throw Error(APRH_ERR_INVALID_ARGUMENT, "stencil sizes must be positive");
    if (sy % 2 == 0 || sx % 2 == 0 || sz % 2 == 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "stencil sizes must be odd");
    if ((int64_t)sy * sx * sz > (1 << 20)) throw Error(APRH_ERR_INVALID_ARGUMENT, "stencil too large");
    HostStencil s;
    s.sy = sy;
    s.sx = sx;
    s.sz = sz;
    s.w.assign(w, w + (size_t)sy * sx * sz);
    return s;
}

std::vector<HostStencil> stencil_list(const float* stencils, const int32_t* dims, int n) {
    if (!stencils || !dims || n < 1) throw Error(APRH_ERR_INVALID_ARGUMENT, "need at least one stencil");
    std::vector<HostStencil> v;
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        v.push_back(host_stencil(stencils + off, dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]));
        off += v.back().w.size();
    }
    return v;
}

int n_stencil_levels(const aprh_apr& a, int use_ds) { return use_ds ? a.l_max - a.l_min : 1; }

/* APRStencil::get_rescaled_stencils (APRStencil.hpp:135-152): the stencil divided by 2^d for
   level deltas d = 0 .. nlevels-1 */
std::vector<HostStencil> rescaled_stencils(const HostStencil& s, int nlevels) {
    std::vector<HostStencil> v((size_t)std::max(nlevels, 1), s);
    for (size_t d = 0; d < v.size(); ++d) {
        const float step = (float)std::pow(2.0, (double)d);
        for (float& w : v[d].w) w = w / step;
    }
    return v;
}

/* central difference along `dim` (APRNumerics.hpp:165-184) */
HostStencil cfd_stencil(int dim, float delta) {
    HostStencil s;
    s.sy = dim == 0 ? 3 : 1;
    s.sx = dim == 1 ? 3 : 1;
    s.sz = dim == 2 ? 3 : 1;
    s.w = {-1.0f / (2 * delta), 0.0f, 1.0f / (2 * delta)};
    return s;
}

/* a 3-tap stencil along one axis as a 3x3x3 stencil: the taps on the centre line, zeros elsewhere */
HostStencil cube3(const HostStencil& s) {
    HostStencil c;
    c.sy = c.sx = c.sz = 3;
    c.w.assign(27, 0.0f);
    for (int k = 0; k < 3; ++k) {
        const int y = s.sy == 3 ? k : 1, x = s.sx == 3 ? k : 1, z = s.sz == 3 ? k : 1;
        c.w[(size_t)(z * 3 + x) * 3 + y] = s.w[(size_t)k];
    }
    return c;
}

/* APRStencil::create_sobel_filter (APRStencil.hpp:254-276) and the 2-D form (:278-297):
   difference along `dim`, [1/4, 1/2, 1/4] smoothing along the others, (y, x, z) products */
HostStencil sobel_stencil(int dim, float delta, bool two_d) {
    const float sm[3] = {0.25f, 0.5f, 0.25f};
    const float df[3] = {-1.0f / (2 * delta), 0.0f, 1.0f / (2 * delta)};
    const float* f[3] = {dim == 0 ? df : sm, dim == 1 ? df : sm, dim == 2 ? df : sm};
    HostStencil s;
    s.sy = 3;
    s.sx = 3;
    s.sz = two_d ? 1 : 3;
    s.w.assign((size_t)(9 * s.sz), 0.0f);
    for (int k = 0; k < s.sz; ++k)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) s.w[(size_t)((k * 3 + j) * 3 + i)] = two_d ? f[0][i] * f[1][j] : f[0][i] * f[1][j] * f[2][k];
    return s;
}

/* the stencil APRNumerics::gradient_cfd / gradient_sobel convolve with (Sobel falls back to the
   central difference below three dimensions, except along y or x of a 2-D image) */
HostStencil gradient_stencil(const aprh_apr& a, int dim, float delta, bool sobel) {
    if (dim < 0 || dim > 2) throw Error(APRH_ERR_INVALID_ARGUMENT, "dimension must be 0 (y), 1 (x) or 2 (z)");
    const int nd = (a.org[0] > 1) + (a.org[1] > 1) + (a.org[2] > 1);
    if (sobel && nd == 3) return sobel_stencil(dim, delta, false);
    if (sobel && nd == 2 && dim != 2) return sobel_stencil(dim, delta, true);
    return cfd_stencil(dim, delta);
}

template <typename T>
void gradient_along(aprh_context* ctx, aprh_apr& a, const T* in, int dim, float delta, bool sobel, float* out) {
    const std::vector<HostStencil> v = rescaled_stencils(gradient_stencil(a, dim, delta, sobel), a.l_max - a.l_min);
    convolve<T>(ctx, a, v, in, out, true);
}

template <typename T>
void gradient_magnitude(aprh_context* ctx, aprh_apr& a, const T* in, const float* deltas, bool sobel, float* out) {
    gradient_along<T>(ctx, a, in, 0, deltas[0], sobel, out);
    Buf gx, gz;
    if (a.org[1] > 1) {
        gx.reset(ctx->pool, a.total * 4);
        gradient_along<T>(ctx, a, in, 1, deltas[1], sobel, gx.as<float>());
    }
    if (a.org[2] > 1) {
        gz.reset(ctx->pool, a.total * 4);
        gradient_along<T>(ctx, a, in, 2, deltas[2], sobel, gz.as<float>());
    }
    gradient_magnitude_combine(ctx, out, a.org[1] > 1 ? gx.as<float>() : nullptr, a.org[2] > 1 ? gz.as<float>() : nullptr, a.total);
}

/* APRNumerics::richardson_lucy_tv (APRNumerics.hpp:489-525): Richardson-Lucy as
   richardson_lucy (the re-blur's quotient in its convolution's epilogue), then div_norm_grad
   (:431-486, unit pixel sizes) of u: central differences along the axes present, normalised by
   their magnitude, and the sum of their differences; u = u * correction / (1 - reg * div) */
template <typename T>
void richardson_lucy_tv(aprh_context* ctx, aprh_apr& a, const T* in, float* u, const std::vector<HostStencil>& psf,
                        const std::vector<HostStencil>& psf_flipped, int niter, float reg, bool resume) {
    const size_t P = a.total;
    if (!resume) fill_f32(ctx, u, P, 1.0f);
    if (P == 0) return;
    ensure_tree(ctx, a);
    const bool sep = ctx->conv_mode == APRH_CONV_SEPARABLE;
    const bool in_u16 = std::is_same<T, uint16_t>::value;
    const bool has_x = a.org[1] > 1, has_z = a.org[2] > 1;
    Buf blur(ctx->pool, P * 4), err(ctx->pool, P * 4), gy(ctx->pool, P * 4), dv(ctx->pool, P * 4), tmp(ctx->pool, P * 4);
    Buf gx, gz, tree(ctx->pool, std::max<uint64_t>(a.t_total, 1) * 4);
    if (has_x) gx.reset(ctx->pool, P * 4);
    if (has_z) gz.reset(ctx->pool, P * 4);
    /* gradient_cfd's rescaled central differences (unit pixel size).  In separable mode each is
       embedded in a 3x3x3 stencil (its taps on the centre line, zeros elsewhere), which the
       register-ring kernel takes; in the reference's tap order they stay 1-D */
    std::vector<HostStencil> cfd[3];
    for (int d = 0; d < 3; ++d) cfd[d] = rescaled_stencils(sep ? cube3(cfd_stencil(d, 1.0f)) : cfd_stencil(d, 1.0f), a.l_max - a.l_min);
    float* tr = tree.as<float>();
    auto grad = [&](const float* src, int d, float* dst) {
        conv_levels<float>(ctx, a, cfd[d], sep, src, tr, dst, true, 0, nullptr, false, nullptr, nullptr);
    };
    for (int it = 0; it < niter; ++it) {
        tree_mean<float>(ctx, a, u, tr);
        conv_levels<float>(ctx, a, psf_flipped, sep, u, tr, blur.as<float>(), true, 1, in, in_u16, nullptr, nullptr);
        /* the gradients of u use the same tree means of u (APRFilter.hpp:1088) */
        grad(u, 0, gy.as<float>());
        if (has_x) grad(u, 1, gx.as<float>());
        if (has_z) grad(u, 2, gz.as<float>());
        tree_mean<float>(ctx, a, blur.as<float>(), tr);
        conv_levels<float>(ctx, a, psf, sep, blur.as<float>(), tr, err.as<float>(), true, 0, nullptr, false, nullptr, nullptr);
        rl_tv_step(ctx, 0, gy.as<float>(), has_x ? gx.as<float>() : nullptr, has_z ? gz.as<float>() : nullptr, 0.0f, P);
        tree_mean<float>(ctx, a, gy.as<float>(), tr);
        grad(gy.as<float>(), 0, dv.as<float>());
        if (has_x) {
            tree_mean<float>(ctx, a, gx.as<float>(), tr);
            grad(gx.as<float>(), 1, tmp.as<float>());
            rl_tv_step(ctx, 1, dv.as<float>(), tmp.as<float>(), nullptr, 0.0f, P);
        }
        if (has_z) {
            tree_mean<float>(ctx, a, gz.as<float>(), tr);
            grad(gz.as<float>(), 2, tmp.as<float>());
            rl_tv_step(ctx, 1, dv.as<float>(), tmp.as<float>(), nullptr, 0.0f, P);
        }
        rl_tv_step(ctx, 2, u, err.as<float>(), dv.as<float>(), reg, P);
    }
}

void check_apr(aprh_context* ctx, aprh_apr* a) {
    check_ctx(ctx);
    if (!a) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL apr");
    if (a->ctx != ctx) throw Error(APRH_ERR_INVALID_ARGUMENT, "apr belongs to another context");
}

}  // namespace

extern "C" {

int aprh_apr_from_structure(aprh_context* ctx, int64_t y, int64_t x, int64_t z, const uint64_t* xz_end_vec, uint64_t xz_end_size,
                            const uint16_t* y_vec, uint64_t total, aprh_apr** out) {
    APRH_TRY
    check_ctx(ctx);
    if (!xz_end_vec || !out || (total > 0 && !y_vec)) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (y <= 0 || x <= 0 || z <= 0) throw Error(APRH_ERR_BAD_DIMS, "dimensions must be positive");
    if (y > 65535 || x > 65535 || z > 65535) throw Error(APRH_ERR_UNSUPPORTED, "dimensions above 65535");
    aprh_apr* a = new aprh_apr();
    a->ctx = ctx;
    try {
        geninfo_init(*a, y, x, z);
        if (xz_end_size != a->n_xz) throw Error(APRH_ERR_INVALID_ARGUMENT, "xz_end_size does not match the dimensions");
        a->total = total;
        a->xz_end.reset(ctx->pool, a->n_xz * 8);
        a->y_vec.reset(ctx->pool, std::max<uint64_t>(total, 1) * 2);
        copy_any(ctx, a->xz_end.ptr(), xz_end_vec, a->n_xz * 8);
        if (total) copy_any(ctx, a->y_vec.ptr(), y_vec, total * 2);
        validate_structure(ctx, *a);
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        delete a;
        throw;
    }
    *out = a;
    APRH_CATCH
}

int aprh_apr_tree_info(aprh_context* ctx, aprh_apr* a, aprh_tree_info* info) {
    APRH_TRY
    check_apr(ctx, a);
    if (!info) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL info");
    ensure_tree(ctx, *a);
    info->l_min = 0;
    info->l_max = a->t_l_max;
    info->total_number_tree_particles = a->t_total;
    info->xz_end_size = a->t_n_xz;
    info->level_xz_size = a->t_level_xz.size();
    APRH_CATCH
}

int aprh_apr_copy_tree(aprh_context* ctx, aprh_apr* a, uint64_t* level_xz_vec, uint64_t* xz_end_vec, uint16_t* y_vec) {
    APRH_TRY
    check_apr(ctx, a);
    ensure_tree(ctx, *a);
    if (level_xz_vec) std::copy(a->t_level_xz.begin(), a->t_level_xz.end(), level_xz_vec);
    if (xz_end_vec) copy_any(ctx, xz_end_vec, a->t_xz_end.ptr(), a->t_n_xz * 8);
    if (y_vec && a->t_total) copy_any(ctx, y_vec, a->t_y_vec.ptr(), a->t_total * 2);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    APRH_CATCH
}

int aprh_fill_tree_mean(aprh_context* ctx, aprh_apr* a, const void* parts, int dtype, float* tree_out) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts || !tree_out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t es = dtype_size(dtype);
    ensure_tree(ctx, *a);
    DevIn in(ctx, parts, std::max<uint64_t>(a->total, 1) * es);
    DevOut out(ctx, tree_out, std::max<uint64_t>(a->t_total, 1) * 4);
    if (dtype == APRH_DTYPE_U16) tree_mean<uint16_t>(ctx, *a, (const uint16_t*)in.ptr, (float*)out.ptr);
    else tree_mean<float>(ctx, *a, (const float*)in.ptr, (float*)out.ptr);
    out.finish(ctx);
    APRH_CATCH
}

int aprh_downsampled_stencils(const float* stencil, int32_t sy, int32_t sx, int32_t sz, int32_t nlevels, int normalize, float* out,
                              int64_t out_capacity, int32_t* dims_out, int64_t* needed) {
    APRH_TRY
    const HostStencil s = host_stencil(stencil, sy, sx, sz);
    if (nlevels < 1) throw Error(APRH_ERR_INVALID_ARGUMENT, "nlevels must be >= 1");
    const std::vector<HostStencil> v = downsampled_stencils(s, nlevels, normalize != 0);
    int64_t tot = 0;
    for (const auto& h : v) tot += (int64_t)h.w.size();
    if (needed) *needed = tot;
    if (out) {
        if (out_capacity < tot) throw Error(APRH_ERR_INVALID_ARGUMENT, "output capacity too small");
        int64_t off = 0;
        for (size_t i = 0; i < v.size(); ++i) {
            std::copy(v[i].w.begin(), v[i].w.end(), out + off);
            off += (int64_t)v[i].w.size();
            if (dims_out) {
                dims_out[3 * i] = v[i].sy;
                dims_out[3 * i + 1] = v[i].sx;
                dims_out[3 * i + 2] = v[i].sz;
            }
        }
    }
    APRH_CATCH
}

int aprh_gaussian_stencil(const float* sigma3, const int32_t* size3, int normalize, float* out) {
    APRH_TRY
    if (!sigma3 || !size3 || !out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int d = 0; d < 3; ++d)
        if (size3[d] < 1 || size3[d] > 4096) throw Error(APRH_ERR_INVALID_ARGUMENT, "stencil size out of range");
    /* APRStencil::create_gaussian_filter (APRStencil.hpp:173-228): separable factors
       expf(-i^2 / (2 s^2)) / (s sqrt(2 pi)), product in (y, x, z) order, float sum over
       y-major loops */
    std::vector<float> g[3];
    for (int d = 0; d < 3; ++d) {
        const int n = size3[d];
        const float s = sigma3[d];
        g[d].assign((size_t)n, 1.0f);
        if (n > 1) {
            int c = 0;
            for (int i = -n / 2; i <= n / 2; ++i) g[d][(size_t)c++] = expf(-i * i / (2 * s * s)) / (s * sqrtf(2 * M_PI));
        }
    }
    const int sy = size3[0], sx = size3[1], sz = size3[2];
    float sum = 0;
    for (int i = 0; i < sy; ++i)
        for (int j = 0; j < sx; ++j)
            for (int k = 0; k < sz; ++k) {
                const float v = g[0][(size_t)i] * g[1][(size_t)j] * g[2][(size_t)k];
                out[((size_t)k * sx + j) * sy + i] = v;
                sum += v;
            }
    if (normalize)
        for (size_t i = 0; i < (size_t)sy * sx * sz; ++i) out[i] /= sum;
    APRH_CATCH
}

int aprh_mean_stencil(const int32_t* size3, float* out) {
    APRH_TRY
    if (!size3 || !out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int d = 0; d < 3; ++d)
        if (size3[d] < 1 || size3[d] > 4096) throw Error(APRH_ERR_INVALID_ARGUMENT, "stencil size out of range");
    /* APRStencil::create_mean_filter (APRStencil.hpp:236-244) */
    const size_t n = (size_t)size3[0] * size3[1] * size3[2];
    const float cnt = (float)n;
    const float v = (float)(1.0 / cnt);
    std::fill(out, out + n, v);
    APRH_CATCH
}

int aprh_convolve_pencil(aprh_context* ctx, aprh_apr* a, const float* stencils, const int32_t* dims, int32_t n_stencils,
                         const void* parts_in, int dtype, float* parts_out, int reflect) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts_in || !parts_out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t es = dtype_size(dtype);
    const std::vector<HostStencil> v = stencil_list(stencils, dims, n_stencils);
    DevIn in(ctx, parts_in, std::max<uint64_t>(a->total, 1) * es);
    DevOut out(ctx, parts_out, std::max<uint64_t>(a->total, 1) * 4);
    if (dtype == APRH_DTYPE_U16) convolve<uint16_t>(ctx, *a, v, (const uint16_t*)in.ptr, (float*)out.ptr, reflect != 0);
    else convolve<float>(ctx, *a, v, (const float*)in.ptr, (float*)out.ptr, reflect != 0);
    out.finish(ctx);
    APRH_CATCH
}

int aprh_convolve(aprh_context* ctx, aprh_apr* a, const float* stencils, const int32_t* dims, int32_t n_stencils, const void* parts_in,
                  int dtype, float* parts_out, int reflect) {
    return aprh_convolve_pencil(ctx, a, stencils, dims, n_stencils, parts_in, dtype, parts_out, reflect);
}

int aprh_convolve_stencil(aprh_context* ctx, aprh_apr* a, const float* stencil, int32_t sy, int32_t sx, int32_t sz, const void* parts_in,
                          int dtype, float* parts_out, int reflect, int use_stencil_downsample, int normalize) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts_in || !parts_out) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t es = dtype_size(dtype);
    const std::vector<HostStencil> v =
        downsampled_stencils(host_stencil(stencil, sy, sx, sz), n_stencil_levels(*a, use_stencil_downsample), normalize != 0);
    DevIn in(ctx, parts_in, std::max<uint64_t>(a->total, 1) * es);
    DevOut out(ctx, parts_out, std::max<uint64_t>(a->total, 1) * 4);
    if (dtype == APRH_DTYPE_U16) convolve<uint16_t>(ctx, *a, v, (const uint16_t*)in.ptr, (float*)out.ptr, reflect != 0);
    else convolve<float>(ctx, *a, v, (const float*)in.ptr, (float*)out.ptr, reflect != 0);
    out.finish(ctx);
    APRH_CATCH
}

int aprh_richardson_lucy(aprh_context* ctx, aprh_apr* a, const void* parts_in, int dtype, float* u, const float* psf, int32_t sy,
                         int32_t sx, int32_t sz, int32_t max_iter, int use_stencil_downsample, int normalize, int resume, double tol,
                         aprh_rl_stats* stats) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts_in || !u) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (max_iter < 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "negative iteration count");
    const size_t es = dtype_size(dtype);
    /* APRNumerics.hpp:575-593: flipped PSF and the two stencil stacks */
    const HostStencil h = host_stencil(psf, sy, sx, sz);
    HostStencil hf = h;
    std::reverse(hf.w.begin(), hf.w.end());
    const int nst = n_stencil_levels(*a, use_stencil_downsample);
    const std::vector<HostStencil> sv = downsampled_stencils(h, nst, normalize != 0);
    const std::vector<HostStencil> fv = downsampled_stencils(hf, nst, normalize != 0);
    DevIn in(ctx, parts_in, std::max<uint64_t>(a->total, 1) * es);
    const bool u_dev = is_device_ptr(u);
    Buf ubuf;
    float* du = u;
    if (!u_dev) {
        ubuf.reset(ctx->pool, std::max<uint64_t>(a->total, 1) * 4);
        du = ubuf.as<float>();
        if (resume && a->total) HIP_CHECK(hipMemcpyAsync(du, u, a->total * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    double rel = 0;
    int it;
    if (dtype == APRH_DTYPE_U16)
        it = richardson_lucy<uint16_t>(ctx, *a, (const uint16_t*)in.ptr, du, sv, fv, max_iter, resume != 0, tol, &rel);
    else
        it = richardson_lucy<float>(ctx, *a, (const float*)in.ptr, du, sv, fv, max_iter, resume != 0, tol, &rel);
    if (!u_dev && a->total) HIP_CHECK(hipMemcpyAsync(u, du, a->total * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (stats) {
        stats->iterations = it;
        stats->last_rel_update = rel;
    }
    APRH_CATCH
}

int aprh_richardson_lucy_tv(aprh_context* ctx, aprh_apr* a, const void* parts_in, int dtype, float* u, const float* psf, int32_t sy,
                            int32_t sx, int32_t sz, int32_t number_iterations, float reg_factor, int use_stencil_downsample, int normalize,
                            int resume) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts_in || !u) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (number_iterations < 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "negative iteration count");
    const size_t es = dtype_size(dtype);
    /* APRNumerics.hpp:549-572: flipped PSF and the two stencil stacks */
    const HostStencil h = host_stencil(psf, sy, sx, sz);
    HostStencil hf = h;
    std::reverse(hf.w.begin(), hf.w.end());
    const int nst = n_stencil_levels(*a, use_stencil_downsample);
    const std::vector<HostStencil> sv = downsampled_stencils(h, nst, normalize != 0);
    const std::vector<HostStencil> fv = downsampled_stencils(hf, nst, normalize != 0);
    DevIn in(ctx, parts_in, std::max<uint64_t>(a->total, 1) * es);
    const bool u_dev = is_device_ptr(u);
    Buf ubuf;
    float* du = u;
    if (!u_dev) {
        ubuf.reset(ctx->pool, std::max<uint64_t>(a->total, 1) * 4);
        du = ubuf.as<float>();
        if (resume && a->total) HIP_CHECK(hipMemcpyAsync(du, u, a->total * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    if (dtype == APRH_DTYPE_U16)
        richardson_lucy_tv<uint16_t>(ctx, *a, (const uint16_t*)in.ptr, du, sv, fv, number_iterations, reg_factor, resume != 0);
    else
        richardson_lucy_tv<float>(ctx, *a, (const float*)in.ptr, du, sv, fv, number_iterations, reg_factor, resume != 0);
    if (!u_dev && a->total) HIP_CHECK(hipMemcpyAsync(u, du, a->total * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    APRH_CATCH
}

int aprh_reconstruct_constant(aprh_context* ctx, aprh_apr* a, const void* parts, int parts_dtype, int level_delta, void* image,
                              int image_dtype) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts || !image) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t pes = dtype_size(parts_dtype), ies = dtype_size(image_dtype);
    if (level_delta > 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "level_delta must be <= 0");
    const int level = a->l_max + level_delta;
    if (level < a->l_min) throw Error(APRH_ERR_INVALID_ARGUMENT, "level_delta below the APR's minimum level");
    const size_t n = (size_t)(a->y_num[level] * a->x_num[level] * a->z_num[level]);
    DevIn in(ctx, parts, std::max<uint64_t>(a->total, 1) * pes);
    DevOut out(ctx, image, n * ies);
    reconstruct_any(ctx, *a, in.ptr, parts_dtype, level, out.ptr, image_dtype);
    out.finish(ctx);
    APRH_CATCH
}

int aprh_reconstruct_level(aprh_context* ctx, aprh_apr* a, void* image, int image_dtype) {
    APRH_TRY
    check_apr(ctx, a);
    if (!image) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t ies = dtype_size(image_dtype);
    const size_t n = (size_t)(a->y_num[a->l_max] * a->x_num[a->l_max] * a->z_num[a->l_max]);
    DevOut out(ctx, image, n * ies);
    reconstruct_level(ctx, *a, a->l_max, out.ptr, image_dtype);
    out.finish(ctx);
    APRH_CATCH
}

int aprh_reconstruct_smooth(aprh_context* ctx, aprh_apr* a, const void* parts, int parts_dtype, const float* scale3, void* image,
                            int image_dtype) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts || !image) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t pes = dtype_size(parts_dtype), ies = dtype_size(image_dtype);
    float sc[3] = {2.0f, 2.0f, 2.0f};
    if (scale3)
        for (int i = 0; i < 3; ++i) sc[i] = scale3[i];
    const size_t n = (size_t)(a->y_num[a->l_max] * a->x_num[a->l_max] * a->z_num[a->l_max]);
    DevIn in(ctx, parts, std::max<uint64_t>(a->total, 1) * pes);
    DevOut out(ctx, image, n * ies);
    reconstruct_smooth(ctx, *a, in.ptr, parts_dtype, sc, out.ptr, image_dtype);
    out.finish(ctx);
    APRH_CATCH
}

static int gradient_entry(aprh_context* ctx, aprh_apr* a, const void* parts, int dtype, int dim, const float* deltas, bool sobel,
                          float* out) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts || !out || (dim < 0 && !deltas)) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (dim >= 0) gradient_stencil(*a, dim, 1.0f, sobel);   /* rejects a bad dimension before any copy */
    const size_t es = dtype_size(dtype);
    DevIn in(ctx, parts, std::max<uint64_t>(a->total, 1) * es);
    DevOut o(ctx, out, std::max<uint64_t>(a->total, 1) * 4);
    if (dim >= 0) {
        if (dtype == APRH_DTYPE_U16) gradient_along<uint16_t>(ctx, *a, (const uint16_t*)in.ptr, dim, deltas[0], sobel, (float*)o.ptr);
        else gradient_along<float>(ctx, *a, (const float*)in.ptr, dim, deltas[0], sobel, (float*)o.ptr);
    } else {
        if (dtype == APRH_DTYPE_U16) gradient_magnitude<uint16_t>(ctx, *a, (const uint16_t*)in.ptr, deltas, sobel, (float*)o.ptr);
        else gradient_magnitude<float>(ctx, *a, (const float*)in.ptr, deltas, sobel, (float*)o.ptr);
    }
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    o.finish(ctx);
    APRH_CATCH
}

int aprh_gradient_cfd(aprh_context* ctx, aprh_apr* a, const void* parts, int dtype, int dimension, float delta, float* out) {
    if (dimension < 0) dimension = 3;   /* reported as out of range */
    return gradient_entry(ctx, a, parts, dtype, dimension, &delta, false, out);
}

int aprh_gradient_sobel(aprh_context* ctx, aprh_apr* a, const void* parts, int dtype, int dimension, float delta, float* out) {
    if (dimension < 0) dimension = 3;
    return gradient_entry(ctx, a, parts, dtype, dimension, &delta, true, out);
}

int aprh_gradient_magnitude_cfd(aprh_context* ctx, aprh_apr* a, const void* parts, int dtype, const float* deltas, float* out) {
    return gradient_entry(ctx, a, parts, dtype, -1, deltas, false, out);
}

int aprh_gradient_magnitude_sobel(aprh_context* ctx, aprh_apr* a, const void* parts, int dtype, const float* deltas, float* out) {
    return gradient_entry(ctx, a, parts, dtype, -1, deltas, true, out);
}

int aprh_face_neighbour_filter(aprh_context* ctx, aprh_apr* a, const void* parts, int dtype, const float* filter3, int dimension,
                               float* out) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts || !out || !filter3) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (dimension < 0 || dimension > 2) throw Error(APRH_ERR_INVALID_ARGUMENT, "dimension must be 0 (y), 1 (x) or 2 (z)");
    const size_t es = dtype_size(dtype);
    DevIn in(ctx, parts, std::max<uint64_t>(a->total, 1) * es);
    DevOut o(ctx, out, std::max<uint64_t>(a->total, 1) * 4);
    if (in.ptr == o.ptr) throw Error(APRH_ERR_INVALID_ARGUMENT, "input and output must be different buffers");
    face_neighbour_filter(ctx, *a, in.ptr, dtype, (float*)o.ptr, filter3, dimension);
    o.finish(ctx);
    APRH_CATCH
}

int aprh_seperable_face_neighbour_filter(aprh_context* ctx, aprh_apr* a, const void* parts, int dtype, const float* filter3,
                                         int32_t repeats, float* out) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts || !out || !filter3) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (repeats < 0) throw Error(APRH_ERR_INVALID_ARGUMENT, "negative repeats");
    const size_t es = dtype_size(dtype);
    const uint64_t P = a->total;
    DevIn in(ctx, parts, std::max<uint64_t>(P, 1) * es);
    DevOut o(ctx, out, std::max<uint64_t>(P, 1) * 4);
    if (P > 0) {
        /* APRNumerics.hpp:286-302: tmp = input (as float); per repeat y, x, z passes */
        Buf tmp(ctx->pool, P * 4), w(ctx->pool, P * 4);
        if (dtype == APRH_DTYPE_U16) local_std_step(ctx, 0, tmp.as<float>(), in.ptr, P);
        else launch_copy(ctx, tmp.ptr(), in.ptr, P * 4);
        for (int r = 0; r < repeats; ++r) {
            face_neighbour_filter(ctx, *a, tmp.ptr(), APRH_DTYPE_F32, w.as<float>(), filter3, 0);
            face_neighbour_filter(ctx, *a, w.ptr(), APRH_DTYPE_F32, tmp.as<float>(), filter3, 1);
            face_neighbour_filter(ctx, *a, tmp.ptr(), APRH_DTYPE_F32, w.as<float>(), filter3, 2);
            launch_copy(ctx, tmp.ptr(), w.ptr(), P * 4);
        }
        launch_copy(ctx, o.ptr, tmp.ptr(), P * 4);
    }
    o.finish(ctx);
    APRH_CATCH
}

int aprh_local_std(aprh_context* ctx, aprh_apr* a, const void* parts, int dtype, const int32_t* size3, float* out) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts || !out || !size3) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t es = dtype_size(dtype);
    /* APRStencil::create_mean_filter (APRStencil.hpp:236-244): 1 / (sy sx sz) from a double */
    std::vector<float> box((size_t)std::max(size3[0], 1) * std::max(size3[1], 1) * std::max(size3[2], 1));
    HostStencil hs = host_stencil(box.data(), size3[0], size3[1], size3[2]);
    std::fill(hs.w.begin(), hs.w.end(), (float)(1.0 / (float)hs.w.size()));
    const std::vector<HostStencil> v = downsampled_stencils(hs, n_stencil_levels(*a, 1), true);
    DevIn in(ctx, parts, std::max<uint64_t>(a->total, 1) * es);
    DevOut o(ctx, out, std::max<uint64_t>(a->total, 1) * 4);
    if (a->total > 0) {
        /* APRNumerics::local_std (APRNumerics.hpp:357-389): a float copy x with its tree means t;
           m = box(x; t); out = box(x^2; t^2); out = sqrtf(max(out - m^2, 0)) */
        Buf x(ctx->pool, a->total * 4), m(ctx->pool, a->total * 4);
        if (dtype == APRH_DTYPE_U16) local_std_step(ctx, 0, x.as<float>(), in.ptr, a->total);
        else launch_copy(ctx, x.ptr(), in.ptr, a->total * 4);
        ensure_tree(ctx, *a);
        Buf t(ctx->pool, std::max<uint64_t>(a->t_total, 1) * 4);
        tree_mean<float>(ctx, *a, x.as<float>(), t.as<float>());
        convolve<float>(ctx, *a, v, x.as<float>(), m.as<float>(), true, t.as<float>());
        local_std_step(ctx, 1, x.as<float>(), nullptr, a->total);
        local_std_step(ctx, 1, t.as<float>(), nullptr, a->t_total);
        convolve<float>(ctx, *a, v, x.as<float>(), (float*)o.ptr, true, t.as<float>());
        local_std_step(ctx, 2, (float*)o.ptr, m.as<float>(), a->total);
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    o.finish(ctx);
    APRH_CATCH
}

int aprh_patch_check_limits(const aprh_apr* a, aprh_patch* patch) {
    APRH_TRY
    if (!a || !patch) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!patch_limits(*a, *patch)) throw Error(APRH_ERR_INVALID_ARGUMENT, "ReconPatch expects begin < end in all dimensions (and extents within int32)");
    APRH_CATCH
}

/* shared checks of the two patch entry points: clamps *p and returns the patch's pixel count */
static size_t patch_setup(aprh_apr* a, const aprh_patch* patch, aprh_patch* p) {
    if (!patch) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    *p = *patch;
    if (p->level_delta < a->l_min - a->l_max) throw Error(APRH_ERR_INVALID_ARGUMENT, "level_delta below the APR's minimum level");
    if (!patch_limits(*a, *p)) throw Error(APRH_ERR_INVALID_ARGUMENT, "ReconPatch expects begin < end in all dimensions");
    return (size_t)(p->y_end - p->y_begin) * (size_t)(p->x_end - p->x_begin) * (size_t)(p->z_end - p->z_begin);
}

int aprh_reconstruct_constant_patch(aprh_context* ctx, aprh_apr* a, const void* parts, int parts_dtype, const aprh_patch* patch,
                                    void* image, int image_dtype) {
    APRH_TRY
    check_apr(ctx, a);
    if (!parts || !image) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t pes = dtype_size(parts_dtype), ies = dtype_size(image_dtype);
    aprh_patch p;
    const size_t n = patch_setup(a, patch, &p);
    DevIn in(ctx, parts, std::max<uint64_t>(a->total, 1) * pes);
    DevOut out(ctx, image, n * ies);
    reconstruct_patch(ctx, *a, in.ptr, parts_dtype, p, out.ptr, image_dtype);
    out.finish(ctx);
    APRH_CATCH
}

int aprh_reconstruct_level_patch(aprh_context* ctx, aprh_apr* a, const aprh_patch* patch, void* image, int image_dtype) {
    APRH_TRY
    check_apr(ctx, a);
    if (!image) throw Error(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t ies = dtype_size(image_dtype);
    aprh_patch p;
    const size_t n = patch_setup(a, patch, &p);
    DevOut out(ctx, image, n * ies);
    reconstruct_patch(ctx, *a, nullptr, 0, p, out.ptr, image_dtype);
    out.finish(ctx);
    APRH_CATCH
}

}  // extern "C"
