/*
 * oracle_capi.cpp -- C entry points of the CPU oracle (liboracle.so).
 *
 * TEST INFRASTRUCTURE ONLY (see apr_oracle.hpp).  The same symbol set is
 * exported by oracle/_ref/libref_*.so, built from the reference's own headers
 * by ref_driver.cpp, so a test can run one check against either library.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   IIR y/x/z, inverse B-spline      reference KAT arrays + _ref bit equality
 *   gradient magnitude + max ds      reference corner KATs + _ref bit equality
 *                                    + integer closed form floor(sqrt(k)/2)
 *   SAT means, padding               reference KAT arrays + _ref bit equality
 *   full get_apr structure, sampling sphere_120 / sphere_diff_dims fixtures
 *                                    (level/x/y/z/pc images) + _ref equality
 */
#include <type_traits>
#include <cstdio>
#include <cstring>
#include <exception>

#include "apr_oracle.hpp"
#include "cnn_oracle.hpp"

using namespace oracle;

namespace {
struct Handle {
    AprResult r;
    int64_t dims[3];
    bool has_tree = false;
    Tree tree;
    const Tree& get_tree() {
        if (!has_tree) {
            build_tree(r.info, r.lin, tree);
            has_tree = true;
        }
        return tree;
    }
};

std::vector<Mesh<float>> unpack_stencils(const float* flat, const int* dims, int n) {
    std::vector<Mesh<float>> v(n);
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        v[i].resize(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]);
        std::copy(flat + off, flat + off + v[i].size(), v[i].v.begin());
        off += v[i].size();
    }
    return v;
}

template <typename T>
Mesh<T> wrap(const T* p, int64_t y, int64_t x, int64_t z) {
    Mesh<T> m;
    m.y_num = y; m.x_num = x; m.z_num = z;
    m.v.assign(p, p + (size_t)(y * x * z));
    return m;
}
template <typename T>
void unwrap(const Mesh<T>& m, T* p) {
    std::memcpy(p, m.v.data(), m.v.size() * sizeof(T));
}
void set_err(char* buf, int len, const char* msg) {
    if (buf && len > 0) {
        std::snprintf(buf, (size_t)len, "%s", msg);
    }
}
}  // namespace

extern "C" {

int orc_version(void) { return 1; }

/* returns the overflow count of the pass (ComputeGradient.hpp:234-244) */
uint64_t orc_bspline_u16(uint16_t* img, int64_t y, int64_t x, int64_t z, int axis, float lambda, float tol) {
    Mesh<uint16_t> m = wrap(img, y, x, z);
    size_t err = 0;
    if (axis == 0) bspline_filt_rec_y(m, lambda, tol, err);
    else if (axis == 1) bspline_filt_rec_x(m, lambda, tol, err);
    else bspline_filt_rec_z(m, lambda, tol, err);
    unwrap(m, img);
    return err;
}

uint64_t orc_bspline_f32(float* img, int64_t y, int64_t x, int64_t z, int axis, float lambda, float tol) {
    Mesh<float> m = wrap(img, y, x, z);
    size_t err = 0;
    if (axis == 0) bspline_filt_rec_y(m, lambda, tol, err);
    else if (axis == 1) bspline_filt_rec_x(m, lambda, tol, err);
    else bspline_filt_rec_z(m, lambda, tol, err);
    unwrap(m, img);
    return err;
}

/* B-spline coefficient bit patterns: out = [b1, b2, norm, k0, minLen, bc1[k0], bc2[k0], bc3[k0], bc4[k0]] */
int orc_bspline_coeffs(float lambda, float tol, int64_t len, float* out, int cap) {
    BsplineCoeffs c = bspline_coeffs(lambda, tol, (size_t)len);
    int need = 5 + 4 * (int)c.k0;
    if (cap < need) return -need;
    out[0] = c.b1; out[1] = c.b2; out[2] = c.norm_factor; out[3] = (float)c.k0; out[4] = (float)c.minLen;
    for (size_t k = 0; k < c.k0; ++k) {
        out[5 + k] = c.bc1[k];
        out[5 + c.k0 + k] = c.bc2[k];
        out[5 + 2 * c.k0 + k] = c.bc3[k];
        out[5 + 3 * c.k0 + k] = c.bc4[k];
    }
    return need;
}

void orc_inv_bspline_f32(float* img, int64_t y, int64_t x, int64_t z, int axis) {
    Mesh<float> m = wrap(img, y, x, z);
    if (axis == 0) inv_bspline_y(m);
    else if (axis == 1) inv_bspline_x(m);
    else inv_bspline_z(m);
    unwrap(m, img);
}

/* grad must hold ceil(y/2)*ceil(x/2)*ceil(z/2) values, pre-initialised by the caller */
void orc_grad_mag_ds_u16(const uint16_t* in, uint16_t* grad, int64_t y, int64_t x, int64_t z, float hx, float hy, float hz) {
    Mesh<uint16_t> m = wrap(in, y, x, z);
    Mesh<uint16_t> g = wrap(grad, (y + 1) / 2, (x + 1) / 2, (z + 1) / 2);
    grad_mag_ds(m, g, hx, hy, hz);
    unwrap(g, grad);
}

void orc_grad_mag_ds_f32(const float* in, float* grad, int64_t y, int64_t x, int64_t z, float hx, float hy, float hz) {
    Mesh<float> m = wrap(in, y, x, z);
    Mesh<float> g = wrap(grad, (y + 1) / 2, (x + 1) / 2, (z + 1) / 2);
    grad_mag_ds(m, g, hx, hy, hz);
    unwrap(g, grad);
}

/* the gradient stage's mean downsample: sum in float, x / 8.0 in double (ComputeGradient.hpp:95-98) */
void orc_downsample_mean_u16_f32(const uint16_t* in, float* out, int64_t y, int64_t x, int64_t z) {
    Mesh<uint16_t> m = wrap(in, y, x, z);
    Mesh<float> o;
    downsample(m, o, [](float a, float b) { return a + b; }, [](float a) { return (float)((double)a / 8.0); });
    unwrap(o, out);
}

/* generic downsample used by tests: op 0 = mean (x/8.0f), 1 = max */
void orc_downsample_f32(const float* in, float* out, int64_t y, int64_t x, int64_t z, int op) {
    Mesh<float> m = wrap(in, y, x, z);
    Mesh<float> o;
    if (op == 0) downsample(m, o, [](float a, float b) { return a + b; }, [](float a) { return a / 8.0f; });
    else downsample(m, o, [](float a, float b) { return std::max(a, b); }, [](float a) { return a; });
    unwrap(o, out);
}

void orc_sat_mean_f32(float* img, int64_t y, int64_t x, int64_t z, int axis, int64_t offset) {
    Mesh<float> m = wrap(img, y, x, z);
    if (axis == 0) sat_mean_y(m, offset);
    else if (axis == 1) sat_mean_x(m, offset);
    else sat_mean_z(m, offset);
    unwrap(m, img);
}

/* out must hold the padded size; pads are clipped like paddPixels */
void orc_pad_f32(const float* in, float* out, int64_t y, int64_t x, int64_t z, int sy, int sx, int sz) {
    Mesh<float> m = wrap(in, y, x, z), o;
    padd(m, o, sy, sx, sz);
    unwrap(o, out);
}

/* lis: in = downsampled smoothed intensities, out = local intensity scale; intensity: out copy */
void orc_local_intensity_scale(float* lis, float* intensity, int64_t y, int64_t x, int64_t z, const aprh_params* par) {
    Mesh<float> m = wrap(lis, y, x, z), it;
    local_intensity_scale(m, it, *par);
    unwrap(m, lis);
    if (intensity) unwrap(it, intensity);
}

void* orc_get_apr_u16(const uint16_t* img, int64_t y, int64_t x, int64_t z, const aprh_params* par, int keep, char* errbuf,
                      int errlen) {
    try {
        Handle* h = new Handle();
        h->dims[0] = y; h->dims[1] = x; h->dims[2] = z;
        Mesh<uint16_t> m = wrap(img, y, x, z);
        get_apr_u16(h->r, m, *par, keep != 0);
        return h;
    } catch (std::exception& e) {
        set_err(errbuf, errlen, e.what());
        return nullptr;
    }
}

void* orc_get_apr_f32(const float* img, int64_t y, int64_t x, int64_t z, const aprh_params* par, int keep, char* errbuf,
                      int errlen) {
    try {
        Handle* h = new Handle();
        h->dims[0] = y; h->dims[1] = x; h->dims[2] = z;
        Mesh<float> m = wrap(img, y, x, z);
        get_apr_f32(h->r, m, *par, keep != 0);
        return h;
    } catch (std::exception& e) {
        set_err(errbuf, errlen, e.what());
        return nullptr;
    }
}

/* APRReconstruction::reconstruct_smooth (APRReconstruction.hpp:270-295); parts dtype 0 u16 /
   1 f32, image dtype 0 u16 / 1 f32 */
void orc_reconstruct_smooth(void* hp, const void* parts, int pdt, const float* scale3, void* out, int idt) {
    Handle* h = (Handle*)hp;
    if (idt == 0) {
        Mesh<uint16_t> img;
        if (pdt == 0) reconstruct_smooth(h->r.info, h->r.lin, (const uint16_t*)parts, scale3, img);
        else reconstruct_smooth(h->r.info, h->r.lin, (const float*)parts, scale3, img);
        std::copy(img.v.begin(), img.v.end(), (uint16_t*)out);
    } else {
        Mesh<float> img;
        if (pdt == 0) reconstruct_smooth(h->r.info, h->r.lin, (const uint16_t*)parts, scale3, img);
        else reconstruct_smooth(h->r.info, h->r.lin, (const float*)parts, scale3, img);
        std::copy(img.v.begin(), img.v.end(), (float*)out);
    }
}

/* one adaptive mean pass of reconstruct_smooth on a float image (axis 0 y, 1 x, 2 z) with a
   caller-supplied level image: calc_sat_adaptive_y/x/z (APRReconstruction.hpp:528-876) */
void orc_sat_adaptive_f32(float* img, const uint16_t* levels, int64_t y, int64_t x, int64_t z, float scale, int offset_max,
                          int d_max, int axis) {
    Mesh<float> m;
    m.resize(y, x, z, 0.0f);
    std::copy(img, img + y * x * z, m.v.begin());
    Mesh<uint16_t> k;
    k.resize(y, x, z, 0);
    std::copy(levels, levels + y * x * z, k.v.begin());
    if (axis == 0) sat_adaptive_y(m, k, scale, offset_max, d_max);
    else sat_adaptive_outer(m, k, scale, offset_max, d_max, axis);
    std::copy(m.v.begin(), m.v.end(), img);
}

/* APRNumerics::face_neighbour_filter / seperable_face_neighbour_filter (repeats >= 0) */
void orc_face_neighbour_filter_f32(void* hp, const float* in, const float* filter3, int dim, int repeats, float* out) {
    Handle* h = (Handle*)hp;
    if (repeats < 0) face_neighbour_filter(h->r.info, h->r.lin, in, out, filter3, dim);
    else seperable_face_neighbour_filter(h->r.info, h->r.lin, in, out, filter3, repeats);
}

/* APRConverterBatch (APRConverterBatch.hpp:111-175) on an image in memory */
void* orc_get_apr_by_block_u16(const uint16_t* img, int64_t y, int64_t x, int64_t z, const aprh_params* par, int64_t z_block_size,
                               int64_t ghost_z, char* errbuf, int errlen) {
    try {
        Handle* h = new Handle();
        h->dims[0] = y; h->dims[1] = x; h->dims[2] = z;
        Mesh<uint16_t> m = wrap(img, y, x, z);
        get_apr_by_block_u16(h->r, m, *par, z_block_size, ghost_z);
        return h;
    } catch (std::exception& e) {
        set_err(errbuf, errlen, e.what());
        return nullptr;
    }
}

/* ParticleData::sample_parts_from_img_blocked (ParticleData.hpp:292-362) */
void orc_sample_image_blocked_u16(void* hp, const uint16_t* img, int64_t z_block_size, int64_t ghost_z, uint16_t* parts) {
    Handle* h = (Handle*)hp;
    Mesh<uint16_t> m = wrap(img, h->dims[0], h->dims[1], h->dims[2]);
    std::vector<uint16_t> p;
    sample_image_blocked(h->r.info, h->r.lin, m, z_block_size, ghost_z, p);
    std::copy(p.begin(), p.end(), parts);
}

/* get_apr_custom_grad_scale (APRConverter.hpp:169-191): structure from given   */
/* downsampled gradient and local intensity scale; intensity = Ip_th reference  */
void* orc_apr_from_grad_lis(const uint16_t* grad, const float* lis, const float* intensity, int64_t y, int64_t x, int64_t z,
                            const aprh_params* par, float bspline_offset) {
    Handle* h = new Handle();
    h->dims[0] = y; h->dims[1] = x; h->dims[2] = z;
    h->r.info.init(y, x, z);
    h->r.bspline_offset = bspline_offset;
    const int64_t yd = (y + 1) / 2, xd = (x + 1) / 2, zd = (z + 1) / 2;
    Mesh<uint16_t> g = wrap(grad, yd, xd, zd);
    Mesh<float> l = wrap(lis, yd, xd, zd);
    Mesh<float> it = wrap(intensity, yd, xd, zd);
    apr_from_grad_lis(h->r, g, l, it, *par, true);
    return h;
}

void orc_apr_info(void* hp, int64_t* out) {
    Handle* h = (Handle*)hp;
    out[0] = h->r.info.l_min;
    out[1] = h->r.info.l_max;
    out[2] = (int64_t)h->r.lin.total;
    out[3] = (int64_t)h->r.lin.xz_end_vec.size();
    out[4] = (int64_t)h->r.lin.level_xz_vec.size();
    std::memcpy(&out[5], &h->r.bspline_offset, sizeof(float));
}

void orc_apr_get(void* hp, uint64_t* level_xz, uint64_t* xz_end, uint16_t* y_vec) {
    Handle* h = (Handle*)hp;
    std::memcpy(level_xz, h->r.lin.level_xz_vec.data(), h->r.lin.level_xz_vec.size() * 8);
    std::memcpy(xz_end, h->r.lin.xz_end_vec.data(), h->r.lin.xz_end_vec.size() * 8);
    std::memcpy(y_vec, h->r.lin.y_vec.data(), h->r.lin.y_vec.size() * 2);
}

/* which: 0 smoothed(u16,N) 1 grad_raw(u16,Nd) 2 lis_raw(f32,Nd) 3 grad(u16,Nd) 4 lis(f32,Nd) 5 levels(f32,Nd) */
int orc_apr_intermediate(void* hp, int which, void* dst) {
    Handle* h = (Handle*)hp;
    switch (which) {
        case 0: if (h->r.is_float) { if (h->r.smoothed_f.size() == 0) return -1; unwrap(h->r.smoothed_f, (float*)dst); return 0; }
                if (h->r.smoothed.size() == 0) return -1; unwrap(h->r.smoothed, (uint16_t*)dst); return 0;
        case 1: if (h->r.is_float) { if (h->r.grad_raw_f.size() == 0) return -1; unwrap(h->r.grad_raw_f, (float*)dst); return 0; }
                if (h->r.grad_raw.size() == 0) return -1; unwrap(h->r.grad_raw, (uint16_t*)dst); return 0;
        case 2: if (h->r.lis_raw.size() == 0) return -1; unwrap(h->r.lis_raw, (float*)dst); return 0;
        case 3: if (h->r.is_float) { if (h->r.grad_f.size() == 0) return -1; unwrap(h->r.grad_f, (float*)dst); return 0; }
                if (h->r.grad.size() == 0) return -1; unwrap(h->r.grad, (uint16_t*)dst); return 0;
        case 4: if (h->r.lis.size() == 0) return -1; unwrap(h->r.lis, (float*)dst); return 0;
        case 5: if (h->r.levels.size() == 0) return -1; unwrap(h->r.levels, (float*)dst); return 0;
        default: return -1;
    }
}

/* final particle-cell tree level l (u8, dims of level l) */
int orc_apr_pct(void* hp, int level, uint8_t* dst) {
    Handle* h = (Handle*)hp;
    if (level < 0 || level >= (int)h->r.pct.size() || h->r.pct[level].size() == 0) return -1;
    unwrap(h->r.pct[level], dst);
    return 0;
}

void orc_sample_image_u16(void* hp, const uint16_t* img, uint16_t* parts) {
    Handle* h = (Handle*)hp;
    Mesh<uint16_t> m = wrap(img, h->dims[0], h->dims[1], h->dims[2]);
    std::vector<uint16_t> p;
    sample_image(h->r.info, h->r.lin, m, p);
    std::memcpy(parts, p.data(), p.size() * 2);
}

void orc_sample_image_f32(void* hp, const float* img, float* parts) {
    Handle* h = (Handle*)hp;
    Mesh<float> m = wrap(img, h->dims[0], h->dims[1], h->dims[2]);
    std::vector<float> p;
    sample_image(h->r.info, h->r.lin, m, p);
    std::memcpy(parts, p.data(), p.size() * 4);
}

void orc_apr_free(void* hp) { delete (Handle*)hp; }

/* APR from given GenInfo dims + LinearAccess arrays (host copies of another build) */
void* orc_apr_from_structure(int64_t y, int64_t x, int64_t z, const uint64_t* xz_end, uint64_t n_xz, const uint16_t* y_vec,
                             uint64_t total) {
    Handle* h = new Handle();
    h->dims[0] = y; h->dims[1] = x; h->dims[2] = z;
    h->r.info.init(y, x, z);
    h->r.lin.init_xz(h->r.info);
    if (h->r.lin.xz_end_vec.size() != n_xz) { delete h; return nullptr; }
    std::copy(xz_end, xz_end + n_xz, h->r.lin.xz_end_vec.begin());
    h->r.lin.y_vec.assign(y_vec, y_vec + total);
    h->r.lin.total = total;
    return h;
}

/* tree (APR::initialize_tree_dense): out = [l_min, l_max, total, n_xz, n_level_xz] */
void orc_tree_info(void* hp, int64_t* out) {
    Handle* h = (Handle*)hp;
    const Tree& t = h->get_tree();
    out[0] = t.info.l_min; out[1] = t.info.l_max; out[2] = (int64_t)t.lin.total;
    out[3] = (int64_t)t.lin.xz_end_vec.size(); out[4] = (int64_t)t.lin.level_xz_vec.size();
}

void orc_tree_get(void* hp, uint64_t* level_xz, uint64_t* xz_end, uint16_t* y_vec) {
    Handle* h = (Handle*)hp;
    const Tree& t = h->get_tree();
    std::copy(t.lin.level_xz_vec.begin(), t.lin.level_xz_vec.end(), level_xz);
    std::copy(t.lin.xz_end_vec.begin(), t.lin.xz_end_vec.end(), xz_end);
    std::copy(t.lin.y_vec.begin(), t.lin.y_vec.end(), y_vec);
}

void orc_fill_tree_mean_f32(void* hp, const float* parts, float* tree) {
    Handle* h = (Handle*)hp;
    std::vector<float> tr;
    fill_tree_mean(h->r.info, h->r.lin, h->get_tree(), parts, tr);
    std::copy(tr.begin(), tr.end(), tree);
}

void orc_fill_tree_mean_u16(void* hp, const uint16_t* parts, float* tree) {
    Handle* h = (Handle*)hp;
    std::vector<float> tr;
    fill_tree_mean(h->r.info, h->r.lin, h->get_tree(), parts, tr);
    std::copy(tr.begin(), tr.end(), tree);
}

/* get_downsampled_stencils: writes the stencils back to back into out and their
   (y, x, z) dims into dims_out; returns the total number of floats */
int64_t orc_downsampled_stencils(const float* st, int sy, int sx, int sz, int nlevels, int normalize, float* out, int64_t cap,
                                 int* dims_out) {
    Mesh<float> m(sy, sx, sz);
    std::copy(st, st + m.size(), m.v.begin());
    std::vector<Mesh<float>> v;
    downsampled_stencils(m, v, nlevels, normalize != 0);
    int64_t tot = 0;
    for (auto& e : v) tot += (int64_t)e.size();
    if (!out || tot > cap) return tot;
    int64_t off = 0;
    for (size_t i = 0; i < v.size(); ++i) {
        std::copy(v[i].v.begin(), v[i].v.end(), out + off);
        off += (int64_t)v[i].size();
        dims_out[3 * i] = (int)v[i].y_num; dims_out[3 * i + 1] = (int)v[i].x_num; dims_out[3 * i + 2] = (int)v[i].z_num;
    }
    return tot;
}

/* APRFilter::convolve_pencil with a stencil stack (stencils[0] at l_max, ...) */
void orc_convolve_pencil_f32(void* hp, const float* stencils, const int* dims, int n, const float* in, float* out, int reflect) {
    Handle* h = (Handle*)hp;
    auto st = unpack_stencils(stencils, dims, n);
    std::vector<float> tr;
    fill_tree_mean(h->r.info, h->r.lin, h->get_tree(), in, tr);
    convolve_pencil(h->r.info, h->r.lin, h->get_tree(), st, in, tr.data(), out, reflect != 0);
}

void orc_convolve_pencil_u16(void* hp, const float* stencils, const int* dims, int n, const uint16_t* in, float* out, int reflect) {
    Handle* h = (Handle*)hp;
    auto st = unpack_stencils(stencils, dims, n);
    std::vector<float> tr;
    fill_tree_mean(h->r.info, h->r.lin, h->get_tree(), in, tr);
    convolve_pencil(h->r.info, h->r.lin, h->get_tree(), st, in, tr.data(), out, reflect != 0);
}

/* APRFilter::convolve computes the same values in the same order as
   convolve_pencil (full-plane instead of pencil buffer, APRFilter.hpp:971-1077) */
void orc_convolve_f32(void* hp, const float* stencils, const int* dims, int n, const float* in, float* out, int reflect) {
    orc_convolve_pencil_f32(hp, stencils, dims, n, in, out, reflect);
}

/* APRNumerics::richardson_lucy; tol > 0 adds the relative-update stopping rule */
int orc_richardson_lucy_u16(void* hp, const uint16_t* in, float* u, const float* psf, int sy, int sx, int sz, int niter, int use_ds,
                            int normalize, int resume, double tol, double* rel) {
    Handle* h = (Handle*)hp;
    Mesh<float> m(sy, sx, sz);
    std::copy(psf, psf + m.size(), m.v.begin());
    return richardson_lucy(h->r.info, h->r.lin, h->get_tree(), in, u, m, niter, use_ds != 0, normalize != 0, resume != 0, tol, rel);
}

int orc_richardson_lucy_f32(void* hp, const float* in, float* u, const float* psf, int sy, int sx, int sz, int niter, int use_ds,
                            int normalize, int resume, double tol, double* rel) {
    Handle* h = (Handle*)hp;
    Mesh<float> m(sy, sx, sz);
    std::copy(psf, psf + m.size(), m.v.begin());
    return richardson_lucy(h->r.info, h->r.lin, h->get_tree(), in, u, m, niter, use_ds != 0, normalize != 0, resume != 0, tol, rel);
}

/* APRNumerics::richardson_lucy_tv (APRNumerics.hpp:489-525, :549-572) */
void orc_richardson_lucy_tv_u16(void* hp, const uint16_t* in, float* u, const float* psf, int sy, int sx, int sz, int niter, float reg,
                                int use_ds, int normalize, int resume) {
    Handle* h = (Handle*)hp;
    Mesh<float> m(sy, sx, sz);
    std::copy(psf, psf + m.size(), m.v.begin());
    richardson_lucy_tv(h->r.info, h->r.lin, h->get_tree(), in, u, m, niter, reg, use_ds != 0, normalize != 0, resume != 0);
}
void orc_richardson_lucy_tv_f32(void* hp, const float* in, float* u, const float* psf, int sy, int sx, int sz, int niter, float reg,
                                int use_ds, int normalize, int resume) {
    Handle* h = (Handle*)hp;
    Mesh<float> m(sy, sx, sz);
    std::copy(psf, psf + m.size(), m.v.begin());
    richardson_lucy_tv(h->r.info, h->r.lin, h->get_tree(), in, u, m, niter, reg, use_ds != 0, normalize != 0, resume != 0);
}

/* APRNumerics::local_std (APRNumerics.hpp:357-389): the box mean (create_mean_filter,
   APRStencil.hpp:236-244, stencils downsampled and normalised) of the float copy of the input
   with its tree means, and of its square with the squared tree means; then
   sqrtf(max(E[x^2] - E[x]^2, 0)) */
extern "C++" {
template <typename PartT>
static void local_std_impl(Handle* h, const PartT* in, const int* size, float* out) {
    const GenInfo& g = h->r.info;
    const Linear& L = h->r.lin;
    const size_t P = (size_t)L.total;
    Mesh<float> box(size[0], size[1], size[2]);
    const float cnt = (float)box.size();
    std::fill(box.v.begin(), box.v.end(), (float)(1.0 / cnt));
    std::vector<Mesh<float>> st;
    downsampled_stencils(box, st, g.l_max - g.l_min, true);
    std::vector<float> x(P), tr, mean(P);
    for (size_t i = 0; i < P; ++i) x[i] = (float)in[i];
    fill_tree_mean(g, L, h->get_tree(), x.data(), tr);
    convolve_pencil(g, L, h->get_tree(), st, x.data(), tr.data(), mean.data(), true);
    for (auto& v : x) v = v * v;
    for (auto& v : tr) v = v * v;
    convolve_pencil(g, L, h->get_tree(), st, x.data(), tr.data(), out, true);
    for (size_t i = 0; i < P; ++i) {
        const float d = out[i] - mean[i] * mean[i];
        out[i] = std::sqrt(std::max(d, 0.0f));
    }
}
}  // extern "C++"

void orc_local_std_u16(void* hp, const uint16_t* in, const int* size, float* out) { local_std_impl((Handle*)hp, in, size, out); }
void orc_local_std_f32(void* hp, const float* in, const int* size, float* out) { local_std_impl((Handle*)hp, in, size, out); }

/* piecewise-constant reconstruction at level l_max + level_delta; dtypes 0 u16, 1 f32 */
void orc_reconstruct(void* hp, const void* parts, int pdt, int level_delta, void* img, int idt) {
    Handle* h = (Handle*)hp;
    const Tree* t = nullptr;
    std::vector<float> tr;
    if (level_delta < 0) {
        t = &h->get_tree();
        if (pdt == 0) fill_tree_mean(h->r.info, h->r.lin, *t, (const uint16_t*)parts, tr);
        else fill_tree_mean(h->r.info, h->r.lin, *t, (const float*)parts, tr);
    }
    auto run = [&](auto pp, auto* out) {
        using ImgT = std::remove_pointer_t<decltype(out)>;
        Mesh<ImgT> m;
        reconstruct_constant(h->r.info, h->r.lin, t, pp, tr.data(), level_delta, m);
        std::copy(m.v.begin(), m.v.end(), out);
    };
    if (pdt == 0 && idt == 0) run((const uint16_t*)parts, (uint16_t*)img);
    else if (pdt == 0) run((const uint16_t*)parts, (float*)img);
    else if (idt == 0) run((const float*)parts, (uint16_t*)img);
    else run((const float*)parts, (float*)img);
}

void orc_gaussian_stencil(const float* sigma, const int* size, int normalize, float* out) {
    Mesh<float> m;
    gaussian_stencil(sigma, size, normalize != 0, m);
    std::copy(m.v.begin(), m.v.end(), out);
}

/* warm-start network: forward of T tiles x[T][(S+8)^3] -> y[T][S^3] */
void orc_cnn_forward(const float* params, const float* x, int T, int S, int emulate_bf16, float* y) {
    cnn_oracle::Net n(params, emulate_bf16 != 0);
    const int64_t in = (int64_t)(S + 8) * (S + 8) * (S + 8), out = (int64_t)S * S * S;
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < T; ++t) {
        cnn_oracle::Acts A;
        cnn_oracle::forward(n, x + t * in, S, A);
        for (int64_t i = 0; i < out; ++i) y[t * out + i] = (float)A.y[(size_t)i];
    }
}

/* mean squared error over the batch and its gradient (N_PARAMS floats) */
double orc_cnn_grad(const float* params, const float* x, const float* tgt, int T, int S, int emulate_bf16, double* grad) {
    cnn_oracle::Net n(params, emulate_bf16 != 0);
    const int64_t in = (int64_t)(S + 8) * (S + 8) * (S + 8), out = (int64_t)S * S * S;
    const double inv_n = 1.0 / (double)(out * T);
    std::vector<double> g(cnn_oracle::N_PARAMS, 0.0);
    double loss = 0;
    for (int t = 0; t < T; ++t) {
        cnn_oracle::Acts A;
        cnn_oracle::forward(n, x + t * in, S, A);
        loss += cnn_oracle::backward(n, x + t * in, tgt + t * out, A, inv_n, g);
    }
    std::copy(g.begin(), g.end(), grad);
    return loss;
}

int orc_cnn_param_count(void) { return cnn_oracle::N_PARAMS; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
