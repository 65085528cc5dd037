/*
 * apr_amd.hpp -- C++ host facade over the C-ABI in aprh.h, mirroring the
 * reference's host interface for the APR hot path so caller code keeps its
 * shape:
 *
 *   APRConverter<uint16_t>::get_apr      algorithm/APRConverter.hpp:34-134
 *   APRConverterBatch::get_apr           algorithm/APRConverterBatch.hpp:33-175
 *   APR (GenInfo + LinearAccess)         data_structures/APR/APR.hpp:18-61
 *   ParticleData<T>::sample_image        data_structures/APR/particles/ParticleData.hpp:122-127
 *   ParticleData::sample_parts_from_img_blocked  data_structures/APR/particles/ParticleData.hpp:140-141, 267-362
 *   APRTreeNumerics::fill_tree_mean      numerics/APRTreeNumerics.hpp:19
 *   APRStencil::*                        numerics/APRStencil.hpp:12-244
 *   APRFilter::convolve(_pencil)         numerics/APRFilter.hpp:67-149
 *   APRNumerics::richardson_lucy(_tv)    numerics/APRNumerics.hpp:126-148
 *   APRNumerics::gradient_* (cfd, sobel, magnitude)  numerics/APRNumerics.hpp:165-283
 *   APRNumerics::local_std               numerics/APRNumerics.hpp:357-389
 *   APRNumerics::(seperable_)face_neighbour_filter  numerics/APRNumerics.hpp:286-355
 *   APRReconstruction::reconstruct_constant  numerics/APRReconstruction.hpp:80, :297
 *   APRReconstruction::reconstruct_level     numerics/APRReconstruction.hpp:131-171
 *   APRReconstruction::reconstruct_smooth    numerics/APRReconstruction.hpp:270-295
 *   ReconPatch and the patch reconstructions numerics/APRReconstruction.hpp:20-66, :186-196, :296-490
 *
 * Error conventions follow the reference: get_apr returns false on bad
 * dimensions (APRConverter.hpp:86-91), u16 smoothing overflow throws
 * std::invalid_argument (ComputeGradient.hpp:409-412), everything else throws
 * std::runtime_error.  Particle and pixel containers live in host memory; the
 * C-ABI also accepts device pointers for callers that keep data on the GPU.
 */
#ifndef APR_AMD_HPP
#define APR_AMD_HPP

#include <algorithm>
#include <iostream>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "aprh.h"

namespace apr_amd {

namespace detail {
inline void check(int code) {
    if (code == APRH_OK) return;
    const std::string msg = std::string("libaprh: ") + aprh_last_error();
    if (code == APRH_ERR_OVERFLOW) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
template <typename T> struct dtype;
template <> struct dtype<uint16_t> { static constexpr int code = APRH_DTYPE_U16; };
template <> struct dtype<float> { static constexpr int code = APRH_DTYPE_F32; };
}  // namespace detail

/* one HIP stream on one device; every call on it is serialised */
class Context {
public:
    explicit Context(int device = 0) { detail::check(aprh_context_create(device, &h_)); }
    ~Context() { aprh_context_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    aprh_context* get() const { return h_; }
    void set_conv_mode(int mode) { detail::check(aprh_context_set_conv_mode(h_, mode)); }
    static Context& default_context() {
        static Context c(0);
        return c;
    }

private:
    aprh_context* h_ = nullptr;
};

/* PixelData<T> (data_structures/Mesh/PixelData.hpp): y fastest */
template <typename T>
struct PixelData {
    int64_t y_num = 0, x_num = 0, z_num = 0;
    std::vector<T> mesh;
    PixelData() {}
    PixelData(int64_t y, int64_t x, int64_t z) { init(y, x, z); }
    void init(int64_t y, int64_t x, int64_t z) {
        y_num = y;
        x_num = x;
        z_num = z;
        mesh.assign((size_t)(y * x * z), T(0));
    }
    T& at(int64_t y, int64_t x, int64_t z) { return mesh[(size_t)((z * x_num + x) * y_num + y)]; }
    const T& at(int64_t y, int64_t x, int64_t z) const { return mesh[(size_t)((z * x_num + x) * y_num + y)]; }
};

/* APRParameters (algorithm/APRParameters.hpp:11-70), numeric fields */
struct APRParameters : aprh_params {
    APRParameters() : aprh_params(aprh_params_default()) {}
};

class APR {
public:
    APR() {}
    ~APR() { reset(); }
    APR(const APR&) = delete;
    APR& operator=(const APR&) = delete;
    APR(APR&& o) noexcept : h_(o.h_), ctx_(o.ctx_), info_(o.info_) { o.h_ = nullptr; }
    APR& operator=(APR&& o) noexcept {
        if (this != &o) {
            reset();
            h_ = o.h_;
            ctx_ = o.ctx_;
            info_ = o.info_;
            o.h_ = nullptr;
        }
        return *this;
    }
    /* APR from stored LinearAccess arrays (as APRFile::read_apr restores them) */
    static APR from_structure(Context& ctx, int64_t y, int64_t x, int64_t z, const std::vector<uint64_t>& xz_end_vec,
                              const std::vector<uint16_t>& y_vec) {
        APR a;
        aprh_apr* h = nullptr;
        detail::check(aprh_apr_from_structure(ctx.get(), y, x, z, xz_end_vec.data(), xz_end_vec.size(),
                                              y_vec.empty() ? nullptr : y_vec.data(), y_vec.size(), &h));
        a.adopt(h, &ctx);
        return a;
    }
    void adopt(aprh_apr* h, Context* ctx) {
        reset();
        h_ = h;
        ctx_ = ctx;
        detail::check(aprh_apr_info_get(h_, &info_));
    }
    void reset() {
        if (h_) aprh_apr_destroy(h_);
        h_ = nullptr;
    }
    aprh_apr* get() const { return h_; }
    Context& context() const { return *ctx_; }
    uint64_t total_number_particles() const { return info_.total_number_particles; }
    int level_min() const { return info_.l_min; }
    int level_max() const { return info_.l_max; }
    int64_t org_dims(int d) const { return info_.org_dims[d]; }
    uint64_t total_number_tree_particles() const {
        aprh_tree_info t;
        detail::check(aprh_apr_tree_info(ctx_->get(), h_, &t));
        return t.total_number_tree_particles;
    }
    /* LinearAccess arrays (LinearAccess.hpp:33-35) copied to the host */
    void linear_access(std::vector<uint64_t>& level_xz_vec, std::vector<uint64_t>& xz_end_vec, std::vector<uint16_t>& y_vec) const {
        level_xz_vec.resize(info_.level_xz_size);
        xz_end_vec.resize(info_.xz_end_size);
        y_vec.resize(info_.total_number_particles);
        detail::check(aprh_apr_copy_structure(h_, level_xz_vec.data(), xz_end_vec.data(), y_vec.empty() ? nullptr : y_vec.data()));
    }

private:
    aprh_apr* h_ = nullptr;
    Context* ctx_ = nullptr;
    aprh_apr_info info_{};
};

template <typename T>
struct ParticleData {
    std::vector<T> data;
    void init(size_t n) { data.assign(n, T(0)); }
    size_t size() const { return data.size(); }
    T& operator[](size_t i) { return data[i]; }
    const T& operator[](size_t i) const { return data[i]; }
    void fill(T v) { std::fill(data.begin(), data.end(), v); }
    /* ParticleData::sample_image (ParticleData.hpp:122-127): uint16 -> uint16, float -> float */
    void sample_image(APR& apr, const PixelData<T>& img) {
        init(apr.total_number_particles());
        if (std::is_same<T, float>::value)
            detail::check(aprh_sample_image_f32(apr.context().get(), apr.get(), (const float*)img.mesh.data(),
                                                data.empty() ? nullptr : (float*)data.data()));
        else
            detail::check(aprh_sample_image_u16(apr.context().get(), apr.get(), (const uint16_t*)img.mesh.data(),
                                                data.empty() ? nullptr : (uint16_t*)data.data()));
    }
    /* sample_parts_from_img_blocked (ParticleData.hpp:140-141, 267-362), from an image in memory
       instead of a TIFF file: per z block the block's mean pyramid with its ghost planes */
    void sample_parts_from_img_blocked(APR& apr, const PixelData<uint16_t>& img, int z_block_size = 256, int ghost_z = 32) {
        static_assert(std::is_same<T, uint16_t>::value, "blocked sampling takes uint16 images into uint16 particles");
        init(apr.total_number_particles());
        detail::check(aprh_sample_image_blocked_u16(apr.context().get(), apr.get(), img.mesh.data(), z_block_size, ghost_z,
                                                    data.empty() ? nullptr : (uint16_t*)data.data()));
    }
};

template <typename ImageType>
class APRConverter {
    static_assert(std::is_same<ImageType, uint16_t>::value || std::is_same<ImageType, float>::value,
                  "this build converts uint16 and float images");

public:
    APRParameters par;
    explicit APRConverter(Context& ctx = Context::default_context()) : ctx_(&ctx) {}
    /* APRConverter::get_apr (APRConverter.hpp:591-600): false on bad dimensions */
    bool get_apr(APR& apr, const PixelData<ImageType>& img) {
        aprh_apr* h = nullptr;
        int rc;
        if (std::is_same<ImageType, float>::value)
            rc = aprh_get_apr_f32(ctx_->get(), (const float*)img.mesh.data(), img.y_num, img.x_num, img.z_num, &par, &h);
        else
            rc = aprh_get_apr_u16(ctx_->get(), (const uint16_t*)img.mesh.data(), img.y_num, img.x_num, img.z_num, &par, &h);
        if (rc == APRH_ERR_BAD_DIMS) return false;
        detail::check(rc);
        apr.adopt(h, ctx_);
        return true;
    }

private:
    Context* ctx_;
};

/* APRConverterBatch<uint16_t> (APRConverterBatch.hpp:33-175): get_apr over z blocks of
   z_block_size planes with ghost_z ghost planes each side, from an image in memory instead of
   a TIFF file; false on bad dimensions */
class APRConverterBatch {
public:
    APRParameters par;
    int z_block_size = 128;
    int ghost_z = 32;
    explicit APRConverterBatch(Context& ctx = Context::default_context()) : ctx_(&ctx) {}
    bool get_apr(APR& apr, const PixelData<uint16_t>& img) {
        aprh_apr* h = nullptr;
        const int rc = aprh_get_apr_by_block_u16(ctx_->get(), img.mesh.data(), img.y_num, img.x_num, img.z_num, &par, z_block_size,
                                                 ghost_z, &h);
        if (rc == APRH_ERR_BAD_DIMS) return false;
        detail::check(rc);
        apr.adopt(h, ctx_);
        return true;
    }

private:
    Context* ctx_;
};

namespace APRStencil {
/* create_gaussian_filter (APRStencil.hpp:173-228); sigma and size in (y, x, z) */
inline PixelData<float> create_gaussian_filter(const std::vector<float>& sigma = {1, 1, 1}, const std::vector<int>& size = {5, 5, 5},
                                               bool normalize = true) {
    PixelData<float> s(size[0], size[1], size[2]);
    const int32_t sz[3] = {size[0], size[1], size[2]};
    detail::check(aprh_gaussian_stencil(sigma.data(), sz, normalize ? 1 : 0, s.mesh.data()));
    return s;
}
/* create_mean_filter (APRStencil.hpp:236-244) */
inline PixelData<float> create_mean_filter(const std::vector<int>& size = {5, 5, 5}) {
    PixelData<float> s(size[0], size[1], size[2]);
    const int32_t sz[3] = {size[0], size[1], size[2]};
    detail::check(aprh_mean_stencil(sz, s.mesh.data()));
    return s;
}
/* get_downsampled_stencils (APRStencil.hpp:81-97) */
inline void get_downsampled_stencils(const PixelData<float>& in, std::vector<PixelData<float>>& out, int nlevels, bool normalize = false) {
    int64_t need = 0;
    detail::check(aprh_downsampled_stencils(in.mesh.data(), (int32_t)in.y_num, (int32_t)in.x_num, (int32_t)in.z_num, nlevels,
                                            normalize ? 1 : 0, nullptr, 0, nullptr, &need));
    std::vector<float> flat((size_t)need);
    std::vector<int32_t> dims(3 * (size_t)nlevels);
    detail::check(aprh_downsampled_stencils(in.mesh.data(), (int32_t)in.y_num, (int32_t)in.x_num, (int32_t)in.z_num, nlevels,
                                            normalize ? 1 : 0, flat.data(), need, dims.data(), &need));
    out.clear();
    size_t off = 0;
    for (int i = 0; i < nlevels; ++i) {
        PixelData<float> s(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]);
        std::copy(flat.begin() + off, flat.begin() + off + s.mesh.size(), s.mesh.begin());
        off += s.mesh.size();
        out.push_back(std::move(s));
    }
}
}  // namespace APRStencil

namespace APRTreeNumerics {
/* fill_tree_mean (APRTreeNumerics.hpp:63-202) */
template <typename T>
void fill_tree_mean(APR& apr, const ParticleData<T>& parts, ParticleData<float>& tree) {
    tree.init(apr.total_number_tree_particles());
    detail::check(aprh_fill_tree_mean(apr.context().get(), apr.get(), parts.data.data(), detail::dtype<T>::code,
                                      tree.data.empty() ? nullptr : tree.data.data()));
}
}  // namespace APRTreeNumerics

namespace APRFilter {
/* convolve_pencil, vector-of-stencils overload (APRFilter.hpp:89-91) */
template <typename In>
void convolve_pencil(APR& apr, const std::vector<PixelData<float>>& stencils, const ParticleData<In>& in, ParticleData<float>& out,
                     bool reflect_boundary = true) {
    std::vector<float> flat;
    std::vector<int32_t> dims;
    for (const auto& s : stencils) {
        flat.insert(flat.end(), s.mesh.begin(), s.mesh.end());
        dims.push_back((int32_t)s.y_num);
        dims.push_back((int32_t)s.x_num);
        dims.push_back((int32_t)s.z_num);
    }
    out.init(apr.total_number_particles());
    detail::check(aprh_convolve_pencil(apr.context().get(), apr.get(), flat.data(), dims.data(), (int32_t)stencils.size(), in.data.data(),
                                       detail::dtype<In>::code, out.data.data(), reflect_boundary ? 1 : 0));
}
/* convolve_pencil with one stencil (APRFilter.hpp:139-149) */
template <typename In>
void convolve_pencil(APR& apr, const PixelData<float>& stencil, const ParticleData<In>& in, ParticleData<float>& out,
                     bool reflect_boundary = true, bool use_stencil_downsample = true, bool normalize = false) {
    out.init(apr.total_number_particles());
    detail::check(aprh_convolve_stencil(apr.context().get(), apr.get(), stencil.mesh.data(), (int32_t)stencil.y_num,
                                        (int32_t)stencil.x_num, (int32_t)stencil.z_num, in.data.data(), detail::dtype<In>::code,
                                        out.data.data(), reflect_boundary ? 1 : 0, use_stencil_downsample ? 1 : 0, normalize ? 1 : 0));
}
/* convolve (APRFilter.hpp:67-69, :110-120): same results as convolve_pencil */
template <typename In>
void convolve(APR& apr, const std::vector<PixelData<float>>& stencils, const ParticleData<In>& in, ParticleData<float>& out,
              bool reflect_boundary = true) {
    convolve_pencil(apr, stencils, in, out, reflect_boundary);
}
template <typename In>
void convolve(APR& apr, const PixelData<float>& stencil, const ParticleData<In>& in, ParticleData<float>& out, bool reflect_boundary = true,
              bool use_stencil_downsample = true, bool normalize = false) {
    convolve_pencil(apr, stencil, in, out, reflect_boundary, use_stencil_downsample, normalize);
}
}  // namespace APRFilter

namespace APRNumerics {
/* richardson_lucy (APRNumerics.hpp:130-134); optional tol > 0 stops on the relative update */
template <typename In>
aprh_rl_stats richardson_lucy(APR& apr, ParticleData<In>& in, ParticleData<float>& out, PixelData<float>& psf, int number_iterations,
                              bool use_stencil_downsample = true, bool normalize = false, bool resume = false, double tol = 0.0) {
    if (!resume) out.init(apr.total_number_particles());
    aprh_rl_stats st{};
    detail::check(aprh_richardson_lucy(apr.context().get(), apr.get(), in.data.data(), detail::dtype<In>::code, out.data.data(),
                                       psf.mesh.data(), (int32_t)psf.y_num, (int32_t)psf.x_num, (int32_t)psf.z_num, number_iterations,
                                       use_stencil_downsample ? 1 : 0, normalize ? 1 : 0, resume ? 1 : 0, tol, &st));
    return st;
}

/* richardson_lucy_tv (APRNumerics.hpp:142-148, :549-572) */
template <typename In>
void richardson_lucy_tv(APR& apr, ParticleData<In>& in, ParticleData<float>& out, PixelData<float>& psf, int number_iterations,
                        float reg_factor, bool use_stencil_downsample, bool normalize, bool resume) {
    if (!resume) out.init(apr.total_number_particles());
    detail::check(aprh_richardson_lucy_tv(apr.context().get(), apr.get(), in.data.data(), detail::dtype<In>::code, out.data.data(),
                                          psf.mesh.data(), (int32_t)psf.y_num, (int32_t)psf.x_num, (int32_t)psf.z_num,
                                          number_iterations, reg_factor, use_stencil_downsample ? 1 : 0, normalize ? 1 : 0,
                                          resume ? 1 : 0));
}

/* gradient_cfd / gradient_sobel (APRNumerics.hpp:165-213); std::invalid_argument on a bad
   dimension, as the reference */
template <typename In>
void gradient_cfd(APR& apr, const ParticleData<In>& in, ParticleData<float>& out, int dimension, float delta = 1.0f) {
    if (dimension < 0 || dimension > 2) throw std::invalid_argument("APRNumerics::gradient_cfd argument 'dimension' must be 0 (y), 1 (x) or 2 (z)");
    out.init(apr.total_number_particles());
    detail::check(aprh_gradient_cfd(apr.context().get(), apr.get(), in.data.data(), detail::dtype<In>::code, dimension, delta, out.data.data()));
}
template <typename In>
void gradient_sobel(APR& apr, const ParticleData<In>& in, ParticleData<float>& out, int dimension, float delta = 1.0f) {
    if (dimension < 0 || dimension > 2) throw std::invalid_argument("APRNumerics::gradient_sobel argument 'dimension' must be 0 (y), 1 (x) or 2 (z)");
    out.init(apr.total_number_particles());
    detail::check(aprh_gradient_sobel(apr.context().get(), apr.get(), in.data.data(), detail::dtype<In>::code, dimension, delta, out.data.data()));
}
/* gradient_magnitude_cfd / _sobel (APRNumerics.hpp:216-283) */
template <typename In>
void gradient_magnitude_cfd(APR& apr, const ParticleData<In>& in, ParticleData<float>& out, const std::vector<float>& deltas = {1.0f, 1.0f, 1.0f}) {
    out.init(apr.total_number_particles());
    detail::check(aprh_gradient_magnitude_cfd(apr.context().get(), apr.get(), in.data.data(), detail::dtype<In>::code, deltas.data(), out.data.data()));
}
template <typename In>
void gradient_magnitude_sobel(APR& apr, const ParticleData<In>& in, ParticleData<float>& out, const std::vector<float>& deltas = {1.0f, 1.0f, 1.0f}) {
    out.init(apr.total_number_particles());
    detail::check(aprh_gradient_magnitude_sobel(apr.context().get(), apr.get(), in.data.data(), detail::dtype<In>::code, deltas.data(), out.data.data()));
}
/* local_std (APRNumerics.hpp:357-389); size in (y, x, z) */
template <typename In>
void local_std(APR& apr, const ParticleData<In>& in, ParticleData<float>& out, const std::vector<int>& size = {3, 3, 3}) {
    out.init(apr.total_number_particles());
    const int32_t sz[3] = {size[0], size[1], size[2]};
    detail::check(aprh_local_std(apr.context().get(), apr.get(), in.data.data(), detail::dtype<In>::code, sz, out.data.data()));
}
/* face_neighbour_filter (APRNumerics.hpp:305-355); filter = {w_minus, w_self, w_plus} */
template <typename In>
void face_neighbour_filter(APR& apr, const ParticleData<In>& in, ParticleData<float>& out, const std::vector<float>& filter,
                           int dimension) {
    if (filter.size() != 3) throw std::invalid_argument("APRNumerics::face_neighbour_filter expects a filter of size 3");
    out.init(apr.total_number_particles());
    detail::check(aprh_face_neighbour_filter(apr.context().get(), apr.get(), in.data.data(), detail::dtype<In>::code, filter.data(),
                                             dimension, out.data.data()));
}
/* seperable_face_neighbour_filter (APRNumerics.hpp:286-302) */
template <typename In>
void seperable_face_neighbour_filter(APR& apr, const ParticleData<In>& in, ParticleData<float>& out, const std::vector<float>& filter,
                                     int repeats = 1) {
    if (filter.size() != 3) throw std::invalid_argument("APRNumerics::seperable_face_neighbour_filter expects a filter of size 3");
    out.init(apr.total_number_particles());
    detail::check(aprh_seperable_face_neighbour_filter(apr.context().get(), apr.get(), in.data.data(), detail::dtype<In>::code,
                                                       filter.data(), repeats, out.data.data()));
}
}  // namespace APRNumerics

/* ReconPatch (APRReconstruction.hpp:20-66) */
struct ReconPatch {
    int x_begin = 0, x_end = -1, y_begin = 0, y_end = -1, z_begin = 0, z_end = -1;
    int level_delta = 0;
    size_t size() const { return (size_t)(z_end - z_begin) * (x_end - x_begin) * (y_end - y_begin); }
    /* clamps to the image at level l_max + level_delta; false (and a message) when a range is empty */
    bool check_limits(APR& apr) {
        aprh_patch p{x_begin, x_end, y_begin, y_end, z_begin, z_end, level_delta};
        if (aprh_patch_check_limits(apr.get(), &p) != APRH_OK) {
            std::wcerr << "ReconPatch expects begin < end in all dimensions" << std::endl;
            return false;
        }
        x_begin = p.x_begin, x_end = p.x_end, y_begin = p.y_begin, y_end = p.y_end, z_begin = p.z_begin, z_end = p.z_end;
        return true;
    }
};

namespace APRReconstruction {
/* reconstruct_constant (APRReconstruction.hpp:80-120); level_delta < 0 as the patch form (:297-394) */
template <typename U, typename T>
void reconstruct_constant(APR& apr, PixelData<U>& img, const ParticleData<T>& parts, int level_delta = 0) {
    const int64_t s = (int64_t)1 << (-level_delta);
    img.init((apr.org_dims(0) + s - 1) / s, (apr.org_dims(1) + s - 1) / s, (apr.org_dims(2) + s - 1) / s);
    detail::check(aprh_reconstruct_constant(apr.context().get(), apr.get(), parts.data.data(), detail::dtype<T>::code, level_delta,
                                            img.mesh.data(), detail::dtype<U>::code));
}

/* reconstruct_level (APRReconstruction.hpp:131-171): each pixel holds the level of its particle */
template <typename U>
void reconstruct_level(APR& apr, PixelData<U>& img) {
    img.init(apr.org_dims(0), apr.org_dims(1), apr.org_dims(2));
    detail::check(aprh_reconstruct_level(apr.context().get(), apr.get(), img.mesh.data(), detail::dtype<U>::code));
}

/* reconstruct_constant over a patch (APRReconstruction.hpp:186-193, :296-394): as the reference,
   an invalid patch prints a message and leaves img untouched */
template <typename U, typename T>
void reconstruct_constant(APR& apr, PixelData<U>& img, const ParticleData<T>& parts, ReconPatch& patch) {
    if (!patch.check_limits(apr)) {
        std::wcerr << "APRReconstruction::reconstruct_constant: invalid patch size - exiting" << std::endl;
        return;
    }
    img.init(patch.y_end - patch.y_begin, patch.x_end - patch.x_begin, patch.z_end - patch.z_begin);
    const aprh_patch p{patch.x_begin, patch.x_end, patch.y_begin, patch.y_end, patch.z_begin, patch.z_end, patch.level_delta};
    detail::check(aprh_reconstruct_constant_patch(apr.context().get(), apr.get(), parts.data.data(), detail::dtype<T>::code, &p,
                                                  img.mesh.data(), detail::dtype<U>::code));
}

/* reconstruct_level over a patch (APRReconstruction.hpp:397-490) */
template <typename U>
void reconstruct_level(APR& apr, PixelData<U>& img, ReconPatch& patch) {
    if (!patch.check_limits(apr)) {
        std::wcerr << "APRReconstruction::reconstruct_level: invalid patch size - exiting" << std::endl;
        return;
    }
    img.init(patch.y_end - patch.y_begin, patch.x_end - patch.x_begin, patch.z_end - patch.z_begin);
    const aprh_patch p{patch.x_begin, patch.x_end, patch.y_begin, patch.y_end, patch.z_begin, patch.z_end, patch.level_delta};
    detail::check(aprh_reconstruct_level_patch(apr.context().get(), apr.get(), &p, img.mesh.data(), detail::dtype<U>::code));
}
/* reconstruct_smooth (APRReconstruction.hpp:270-295); scale_d in (y, x, z) */
template <typename U, typename T>
void reconstruct_smooth(APR& apr, PixelData<U>& img, const ParticleData<T>& parts, const std::vector<float>& scale_d = {2, 2, 2}) {
    if (scale_d.size() != 3) throw std::invalid_argument("APRReconstruction::reconstruct_smooth expects scale_d of size 3");
    img.init(apr.org_dims(0), apr.org_dims(1), apr.org_dims(2));
    detail::check(aprh_reconstruct_smooth(apr.context().get(), apr.get(), parts.data.data(), detail::dtype<T>::code, scale_d.data(),
                                          img.mesh.data(), detail::dtype<U>::code));
}
}  // namespace APRReconstruction

}  // namespace apr_amd

#endif /* APR_AMD_HPP */
