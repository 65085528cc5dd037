/*
 * ref_driver.cpp -- exports the oracle's C entry points, implemented by the
 * reference's own header-only code (read in place from the reference tree,
 * never copied).  Built only where the reference tree exists, into
 * oracle/_ref/, by oracle/Makefile.  TEST INFRASTRUCTURE ONLY.
 *
 * APRConverter.hpp itself cannot be compiled here: its include chain
 * (APRReconstruction.hpp -> LazyData.hpp -> io/APRFile.hpp) needs HDF5 and
 * the CMake-generated ConfigAPR.h.  The stage classes it drives do compile,
 * so get_apr below calls them in the order of APRConverter::get_apr_cpu
 * (APRConverter.hpp:534-582, computeL :194-247, applyParameters :249-297,
 * solveForAPR :300-336, generateDatastructures :338-363).
 */
#include <cstdio>
#include <cstring>
#include <exception>
#include <stdexcept>

#include "algorithm/APRParameters.hpp"
#include "algorithm/AutoParameters.hpp"
#include "algorithm/ComputeGradient.hpp"
#include "algorithm/LocalIntensityScale.hpp"
#include "algorithm/PullingScheme.hpp"
#include "algorithm/PullingSchemeSparse.hpp"
#include "algorithm/LocalParticleCellSet.hpp"
#include "data_structures/APR/APR.hpp"
#include "data_structures/APR/access/LinearAccess.hpp"
#include "data_structures/APR/particles/ParticleData.hpp"
#include "numerics/APRStencil.hpp"
#include "numerics/APRTreeNumerics.hpp"
#include "numerics/APRFilter.hpp"
#include "numerics/APRNumerics.hpp"

#include "../include/aprh_params.h"

namespace {

APRParameters to_par(const aprh_params* p) {
    APRParameters r;
    r.dx = p->dx; r.dy = p->dy; r.dz = p->dz;
    r.psfx = p->psfx; r.psfy = p->psfy; r.psfz = p->psfz;
    r.Ip_th = p->Ip_th; r.lambda = p->lambda; r.rel_error = p->rel_error;
    r.sigma_th = p->sigma_th; r.sigma_th_max = p->sigma_th_max; r.grad_th = p->grad_th;
    r.reflect_bc_lis = p->reflect_bc_lis != 0;
    r.neighborhood_optimization = p->neighborhood_optimization != 0;
    r.constant_intensity_scale = p->constant_intensity_scale != 0;
    r.auto_parameters = false;
    return r;
}

template <typename T>
void load(PixelData<T>& m, const T* p, int64_t y, int64_t x, int64_t z) {
    m.initWithValue((int)y, (int)x, (int)z, 0);
    std::copy(p, p + (size_t)(y * x * z), m.mesh.begin());
}
template <typename T>
void store(const PixelData<T>& m, T* p) {
    std::copy(m.mesh.begin(), m.mesh.begin() + m.mesh.size(), p);
}

struct Handle {
    APR apr;
    float bspline_offset = 0;
    int64_t dims[3];
    PixelData<uint16_t> smoothed, grad_raw, grad;
    PixelData<float> lis_raw, lis, levels;
    bool is_float = false;                      /* APRConverter<float>: these three instead */
    PixelData<float> smoothed_f, grad_raw_f, grad_f;
    std::vector<PixelData<uint8_t>> pct;
};

/* applyParameters + solveForAPR + generateDatastructures, with the stage    */
/* classes of the reference (APRConverter.hpp:249-363).                      */
void keep(PixelData<uint16_t>& dst16, PixelData<float>&, PixelData<uint16_t>& src) { dst16.init(src); dst16.copyFromMesh(src); }
void keep(PixelData<uint16_t>&, PixelData<float>& dstf, PixelData<float>& src) { dstf.init(src); dstf.copyFromMesh(src); }

template <typename G>
void finish(Handle* h, APRParameters& par, PixelData<G>& grad_temp, PixelData<float>& local_scale_temp,
            PixelData<float>& local_scale_temp2) {
    ComputeGradient cg;
    LocalParticleCellSet lpcs;
    PullingScheme ps;
    par.validate_parameters();
    cg.threshold_gradient(grad_temp, local_scale_temp2, par.Ip_th + h->bspline_offset);
    float max_th = 60000;
    for (size_t i = 0; i < grad_temp.mesh.size(); ++i) {
        float rescaled = local_scale_temp.mesh[i];
        if (rescaled < par.sigma_th) {
            rescaled = (rescaled < par.sigma_th_max) ? max_th : par.sigma_th;
            local_scale_temp.mesh[i] = rescaled;
        }
    }
    for (size_t i = 0; i < grad_temp.mesh.size(); ++i) {
        if (grad_temp.mesh[i] < par.grad_th) grad_temp.mesh[i] = 0;
    }
    keep(h->grad, h->grad_f, grad_temp);
    h->lis.init(local_scale_temp); h->lis.copyFromMesh(local_scale_temp);
    lpcs.computeLevels(grad_temp, local_scale_temp, h->apr.level_max(), par.rel_error, par.dx, par.dy, par.dz);
    h->levels.init(local_scale_temp); h->levels.copyFromMesh(local_scale_temp);
    ps.initialize_particle_cell_tree(h->apr.aprInfo);
    lpcs.get_local_particle_cell_set(ps, local_scale_temp, local_scale_temp2, par);
    ps.pulling_scheme_main();
    h->apr.linearAccess.initialize_linear_structure(h->apr.parameters, ps.getParticleCellTree());
    h->apr.apr_initialized = true;
    auto& t = ps.getParticleCellTree();
    h->pct.resize(t.size());
    for (size_t l = 0; l < t.size(); ++l) {
        if (t[l].mesh.size()) { h->pct[l].init(t[l]); h->pct[l].copyFromMesh(t[l]); }
    }
}

void set_err(char* buf, int len, const char* msg) {
    if (buf && len > 0) std::snprintf(buf, (size_t)len, "%s", msg);
}

template <typename T>
void fill_parts(ParticleData<T>& p, const T* v, size_t n) {
    p.init(n);
    std::copy(v, v + n, p.data.begin());
}

static std::vector<PixelData<float>> unpack(const float* flat, const int* dims, int n) {
    std::vector<PixelData<float>> v(n);
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        v[i].init(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]);
        std::copy(flat + off, flat + off + v[i].mesh.size(), v[i].mesh.begin());
        off += v[i].mesh.size();
    }
    return v;
}

/* initPipelineAPR + computeL + the finishing stages for either image type */
template <typename T>
void* get_apr_ref(const T* img, int64_t y, int64_t x, int64_t z, const aprh_params* p, char* errbuf, int errlen) {
    Handle* h = new Handle();
    try {
        APRParameters par = to_par(p);
        h->dims[0] = y; h->dims[1] = x; h->dims[2] = z;
        h->is_float = std::is_floating_point<T>::value;
        PixelData<T> input;
        load(input, img, y, x, z);
        /* initPipelineAPR (APRConverter.hpp:84-100) */
        h->apr.parameters = par;
        h->apr.aprInfo.init(input.y_num, input.x_num, input.z_num);
        h->apr.linearAccess.genInfo = &h->apr.aprInfo;
        /* initPipelineMemory (APRConverter.hpp:137-167): grad_temp has the image type */
        const int zd = ceil(1.0 * z / 2.0), xd = ceil(1.0 * x / 2.0), yd = ceil(1.0 * y / 2.0);
        PixelData<T> grad_temp;
        PixelData<float> local_scale_temp, local_scale_temp2;
        grad_temp.initWithResize(yd, xd, zd);
        grad_temp.fill(0);
        local_scale_temp.initWithResize(yd, xd, zd);
        local_scale_temp2.initWithResize(yd, xd, zd);
        /* computeL (APRConverter.hpp:194-247) */
        PixelData<T> image_temp(input, false, false);
        if (std::is_floating_point<T>::value) {
            h->bspline_offset = 0;
            image_temp.copyFromMesh(input);
        } else {
            h->bspline_offset = compute_bspline_offset<T>(input, par.lambda);
            const float off = h->bspline_offset;
            image_temp.copyFromMeshWithUnaryOp(input, [=](const auto& a) { return (a + off); });
        }
        ComputeGradient cg;
        /* get_gradient split open to keep the smoothed image (ComputeGradient.hpp:70-119) */
        if (par.lambda > 0) cg.get_smooth_bspline_3D(image_temp, par.lambda);
        keep(h->smoothed, h->smoothed_f, image_temp);
        cg.calc_bspline_fd_ds_mag(image_temp, grad_temp, par.dx, par.dy, par.dz);
        downsample(image_temp, local_scale_temp, [](const float& a, const float& b) -> float { return a + b; },
                   [](const float& a) -> float { return a / 8.0; });
        if (par.lambda > 0) {
            if (image_temp.y_num > 2) cg.calc_inv_bspline_y(local_scale_temp);
            if (image_temp.x_num > 2) cg.calc_inv_bspline_x(local_scale_temp);
            if (image_temp.z_num > 2) cg.calc_inv_bspline_z(local_scale_temp);
        }
        keep(h->grad_raw, h->grad_raw_f, grad_temp);
        LocalIntensityScale lis;
        lis.get_local_intensity_scale(local_scale_temp, local_scale_temp2, par);
        h->lis_raw.init(local_scale_temp); h->lis_raw.copyFromMesh(local_scale_temp);
        finish(h, par, grad_temp, local_scale_temp, local_scale_temp2);
        return h;
    } catch (std::exception& e) {
        set_err(errbuf, errlen, e.what());
        delete h;
        return nullptr;
    }
}

}  // namespace

extern "C" {

int orc_version(void) { return 1000; }

uint64_t orc_bspline_u16(uint16_t* img, int64_t y, int64_t x, int64_t z, int axis, float lambda, float tol) {
    PixelData<uint16_t> m;
    load(m, img, y, x, z);
    ComputeGradient cg;
    uint64_t err = 0;
    try {
        if (axis == 0) cg.bspline_filt_rec_y(m, lambda, tol);
        else if (axis == 1) cg.bspline_filt_rec_x(m, lambda, tol);
        else cg.bspline_filt_rec_z(m, lambda, tol);
    } catch (std::invalid_argument&) {
        err = 1;
    }
    store(m, img);
    return err;
}

uint64_t orc_bspline_f32(float* img, int64_t y, int64_t x, int64_t z, int axis, float lambda, float tol) {
    PixelData<float> m;
    load(m, img, y, x, z);
    ComputeGradient cg;
    if (axis == 0) cg.bspline_filt_rec_y(m, lambda, tol);
    else if (axis == 1) cg.bspline_filt_rec_x(m, lambda, tol);
    else cg.bspline_filt_rec_z(m, lambda, tol);
    store(m, img);
    return 0;
}

int orc_bspline_coeffs(float, float, int64_t, float*, int) { return 0; /* not observable in the reference */ }

void orc_inv_bspline_f32(float* img, int64_t y, int64_t x, int64_t z, int axis) {
    PixelData<float> m;
    load(m, img, y, x, z);
    ComputeGradient cg;
    if (axis == 0) cg.calc_inv_bspline_y(m);
    else if (axis == 1) cg.calc_inv_bspline_x(m);
    else cg.calc_inv_bspline_z(m);
    store(m, img);
}

void orc_grad_mag_ds_u16(const uint16_t* in, uint16_t* grad, int64_t y, int64_t x, int64_t z, float hx, float hy, float hz) {
    PixelData<uint16_t> m, g;
    load(m, in, y, x, z);
    load(g, grad, (y + 1) / 2, (x + 1) / 2, (z + 1) / 2);
    ComputeGradient cg;
    cg.calc_bspline_fd_ds_mag(m, g, hx, hy, hz);
    store(g, grad);
}

void orc_grad_mag_ds_f32(const float* in, float* grad, int64_t y, int64_t x, int64_t z, float hx, float hy, float hz) {
    PixelData<float> m, g;
    load(m, in, y, x, z);
    load(g, grad, (y + 1) / 2, (x + 1) / 2, (z + 1) / 2);
    ComputeGradient cg;
    cg.calc_bspline_fd_ds_mag(m, g, hx, hy, hz);
    store(g, grad);
}

void orc_downsample_mean_u16_f32(const uint16_t* in, float* out, int64_t y, int64_t x, int64_t z) {
    PixelData<uint16_t> m;
    load(m, in, y, x, z);
    PixelData<float> o((int)((y + 1) / 2), (int)((x + 1) / 2), (int)((z + 1) / 2));
    /* same operators as ComputeGradient::get_gradient (ComputeGradient.hpp:95-98) */
    downsample(m, o, [](const float& a, const float& b) -> float { return a + b; },
               [](const float& a) -> float { return a / 8.0; });
    store(o, out);
}

void orc_downsample_f32(const float* in, float* out, int64_t y, int64_t x, int64_t z, int op) {
    PixelData<float> m;
    load(m, in, y, x, z);
    PixelData<float> o;
    if (op == 0) {
        downsample(m, o, [](const float x, const float y) -> float { return x + y; }, [](const float x) -> float { return x / 8.0f; },
                   true);
    } else {
        downsample(m, o, [](const float& a, const float& b) -> float { return std::max(a, b); },
                   [](const float& a) -> float { return a; }, true);
    }
    store(o, out);
}

void orc_sat_mean_f32(float* img, int64_t y, int64_t x, int64_t z, int axis, int64_t offset) {
    PixelData<float> m;
    load(m, img, y, x, z);
    LocalIntensityScale lis;
    if (axis == 0) lis.calc_sat_mean_y(m, (size_t)offset);
    else if (axis == 1) lis.calc_sat_mean_x(m, (size_t)offset);
    else lis.calc_sat_mean_z(m, (size_t)offset);
    store(m, img);
}

void orc_pad_f32(const float* in, float* out, int64_t y, int64_t x, int64_t z, int sy, int sx, int sz) {
    PixelData<float> m, o;
    load(m, in, y, x, z);
    paddPixels(m, o, sy, sx, sz);
    store(o, out);
}

void orc_local_intensity_scale(float* lis, float* intensity, int64_t y, int64_t x, int64_t z, const aprh_params* p) {
    APRParameters par = to_par(p);
    PixelData<float> a, b;
    load(a, lis, y, x, z);
    b.initWithValue((int)y, (int)x, (int)z, 0);
    LocalIntensityScale l;
    l.get_local_intensity_scale(a, b, par);
    store(a, lis);
    if (intensity) store(b, intensity);
}

void* orc_get_apr_u16(const uint16_t* img, int64_t y, int64_t x, int64_t z, const aprh_params* p, int keep_, char* errbuf,
                      int errlen) {
    (void)keep_;
    return get_apr_ref(img, y, x, z, p, errbuf, errlen);
}

void* orc_get_apr_f32(const float* img, int64_t y, int64_t x, int64_t z, const aprh_params* p, int keep_, char* errbuf,
                      int errlen) {
    (void)keep_;
    return get_apr_ref(img, y, x, z, p, errbuf, errlen);
}

void* orc_apr_from_grad_lis(const uint16_t* grad, const float* lisp, const float* intensity, int64_t y, int64_t x, int64_t z,
                            const aprh_params* p, float bspline_offset) {
    Handle* h = new Handle();
    APRParameters par = to_par(p);
    h->dims[0] = y; h->dims[1] = x; h->dims[2] = z;
    h->apr.parameters = par;
    h->apr.aprInfo.init(y, x, z);
    h->apr.linearAccess.genInfo = &h->apr.aprInfo;
    h->bspline_offset = bspline_offset;
    PixelData<uint16_t> g;
    PixelData<float> l, it;
    load(g, grad, (y + 1) / 2, (x + 1) / 2, (z + 1) / 2);
    load(l, lisp, (y + 1) / 2, (x + 1) / 2, (z + 1) / 2);
    load(it, intensity, (y + 1) / 2, (x + 1) / 2, (z + 1) / 2);
    finish(h, par, g, l, it);
    return h;
}

void orc_apr_info(void* hp, int64_t* out) {
    Handle* h = (Handle*)hp;
    out[0] = h->apr.level_min();
    out[1] = h->apr.level_max();
    out[2] = (int64_t)h->apr.linearAccess.y_vec.size();
    out[3] = (int64_t)h->apr.linearAccess.xz_end_vec.size();
    out[4] = (int64_t)h->apr.linearAccess.level_xz_vec.size();
    std::memcpy(&out[5], &h->bspline_offset, sizeof(float));
}

void orc_apr_get(void* hp, uint64_t* level_xz, uint64_t* xz_end, uint16_t* y_vec) {
    Handle* h = (Handle*)hp;
    auto& la = h->apr.linearAccess;
    std::copy(la.level_xz_vec.begin(), la.level_xz_vec.end(), level_xz);
    std::copy(la.xz_end_vec.begin(), la.xz_end_vec.end(), xz_end);
    std::copy(la.y_vec.begin(), la.y_vec.end(), y_vec);
}

int orc_apr_intermediate(void* hp, int which, void* dst) {
    Handle* h = (Handle*)hp;
    switch (which) {
        case 0: if (h->is_float) { store(h->smoothed_f, (float*)dst); return 0; }
                if (!h->smoothed.mesh.size()) return -1; store(h->smoothed, (uint16_t*)dst); return 0;
        case 1: if (h->is_float) { store(h->grad_raw_f, (float*)dst); return 0; }
                if (!h->grad_raw.mesh.size()) return -1; store(h->grad_raw, (uint16_t*)dst); return 0;
        case 2: if (!h->lis_raw.mesh.size()) return -1; store(h->lis_raw, (float*)dst); return 0;
        case 3: if (h->is_float) { store(h->grad_f, (float*)dst); return 0; }
                store(h->grad, (uint16_t*)dst); return 0;
        case 4: store(h->lis, (float*)dst); return 0;
        case 5: store(h->levels, (float*)dst); return 0;
        default: return -1;
    }
}

int orc_apr_pct(void* hp, int level, uint8_t* dst) {
    Handle* h = (Handle*)hp;
    if (level < 0 || level >= (int)h->pct.size() || !h->pct[level].mesh.size()) return -1;
    store(h->pct[level], dst);
    return 0;
}

void orc_sample_image_u16(void* hp, const uint16_t* img, uint16_t* parts) {
    Handle* h = (Handle*)hp;
    PixelData<uint16_t> m;
    load(m, img, h->dims[0], h->dims[1], h->dims[2]);
    ParticleData<uint16_t> p;
    p.sample_image(h->apr, m);
    std::copy(p.data.begin(), p.data.end(), parts);
}

void orc_apr_free(void* hp) { delete (Handle*)hp; }

void* orc_apr_from_structure(int64_t y, int64_t x, int64_t z, const uint64_t* xz_end, uint64_t n_xz, const uint16_t* y_vec,
                             uint64_t total) {
    Handle* h = new Handle();
    h->dims[0] = y; h->dims[1] = x; h->dims[2] = z;
    h->apr.aprInfo.init(y, x, z);
    h->apr.linearAccess.genInfo = &h->apr.aprInfo;
    h->apr.linearAccess.initialize_xz_linear();
    if (h->apr.linearAccess.xz_end_vec.size() != n_xz) { delete h; return nullptr; }
    std::copy(xz_end, xz_end + n_xz, h->apr.linearAccess.xz_end_vec.begin());
    h->apr.linearAccess.y_vec.resize(total);
    std::copy(y_vec, y_vec + total, h->apr.linearAccess.y_vec.begin());
    h->apr.aprInfo.total_number_particles = total;
    h->apr.apr_initialized = true;
    return h;
}

void orc_tree_info(void* hp, int64_t* out) {
    Handle* h = (Handle*)hp;
    auto it = h->apr.tree_iterator();
    out[0] = h->apr.treeInfo.l_min; out[1] = h->apr.treeInfo.l_max;
    out[2] = (int64_t)h->apr.linearAccessTree.y_vec.size();
    out[3] = (int64_t)h->apr.linearAccessTree.xz_end_vec.size();
    out[4] = (int64_t)h->apr.linearAccessTree.level_xz_vec.size();
    (void)it;
}

void orc_tree_get(void* hp, uint64_t* level_xz, uint64_t* xz_end, uint16_t* y_vec) {
    Handle* h = (Handle*)hp;
    auto it = h->apr.tree_iterator();
    (void)it;
    auto& la = h->apr.linearAccessTree;
    std::copy(la.level_xz_vec.begin(), la.level_xz_vec.end(), level_xz);
    std::copy(la.xz_end_vec.begin(), la.xz_end_vec.end(), xz_end);
    std::copy(la.y_vec.begin(), la.y_vec.end(), y_vec);
}

void orc_fill_tree_mean_f32(void* hp, const float* parts, float* tree) {
    Handle* h = (Handle*)hp;
    ParticleData<float> p, t;
    fill_parts(p, parts, h->apr.total_number_particles());
    APRTreeNumerics::fill_tree_mean(h->apr, p, t);
    std::copy(t.data.begin(), t.data.end(), tree);
}

void orc_fill_tree_mean_u16(void* hp, const uint16_t* parts, float* tree) {
    Handle* h = (Handle*)hp;
    ParticleData<uint16_t> p;
    ParticleData<float> t;
    fill_parts(p, parts, h->apr.total_number_particles());
    APRTreeNumerics::fill_tree_mean(h->apr, p, t);
    std::copy(t.data.begin(), t.data.end(), tree);
}

int64_t orc_downsampled_stencils(const float* st, int sy, int sx, int sz, int nlevels, int normalize, float* out, int64_t cap,
                                 int* dims_out) {
    PixelData<float> m(sy, sx, sz);
    std::copy(st, st + (size_t)sy * sx * sz, m.mesh.begin());
    std::vector<PixelData<float>> v;
    APRStencil::get_downsampled_stencils(m, v, nlevels, normalize != 0);
    int64_t tot = 0;
    for (auto& e : v) tot += (int64_t)e.mesh.size();
    if (!out || tot > cap) return tot;
    int64_t off = 0;
    for (size_t i = 0; i < v.size(); ++i) {
        std::copy(v[i].mesh.begin(), v[i].mesh.begin() + v[i].mesh.size(), out + off);
        off += (int64_t)v[i].mesh.size();
        dims_out[3 * i] = v[i].y_num; dims_out[3 * i + 1] = v[i].x_num; dims_out[3 * i + 2] = v[i].z_num;
    }
    return tot;
}

void orc_convolve_pencil_f32(void* hp, const float* stencils, const int* dims, int n, const float* in, float* out, int reflect) {
    Handle* h = (Handle*)hp;
    auto st = unpack(stencils, dims, n);
    ParticleData<float> pi, po;
    fill_parts(pi, in, h->apr.total_number_particles());
    APRFilter::convolve_pencil(h->apr, st, pi, po, reflect != 0);
    std::copy(po.data.begin(), po.data.end(), out);
}

void orc_convolve_pencil_u16(void* hp, const float* stencils, const int* dims, int n, const uint16_t* in, float* out, int reflect) {
    Handle* h = (Handle*)hp;
    auto st = unpack(stencils, dims, n);
    ParticleData<uint16_t> pi;
    ParticleData<float> po;
    fill_parts(pi, in, h->apr.total_number_particles());
    APRFilter::convolve_pencil(h->apr, st, pi, po, reflect != 0);
    std::copy(po.data.begin(), po.data.end(), out);
}

void orc_convolve_f32(void* hp, const float* stencils, const int* dims, int n, const float* in, float* out, int reflect) {
    Handle* h = (Handle*)hp;
    auto st = unpack(stencils, dims, n);
    ParticleData<float> pi, po;
    fill_parts(pi, in, h->apr.total_number_particles());
    APRFilter::convolve(h->apr, st, pi, po, reflect != 0);
    std::copy(po.data.begin(), po.data.end(), out);
}

void orc_local_std_u16(void* hp, const uint16_t* in, const int* size, float* out) {
    Handle* h = (Handle*)hp;
    ParticleData<uint16_t> pi;
    ParticleData<float> po;
    fill_parts(pi, in, h->apr.total_number_particles());
    APRNumerics::local_std(h->apr, pi, po, std::vector<int>{size[0], size[1], size[2]});
    std::copy(po.data.begin(), po.data.end(), out);
}

void orc_local_std_f32(void* hp, const float* in, const int* size, float* out) {
    Handle* h = (Handle*)hp;
    ParticleData<float> pi, po;
    fill_parts(pi, in, h->apr.total_number_particles());
    APRNumerics::local_std(h->apr, pi, po, std::vector<int>{size[0], size[1], size[2]});
    std::copy(po.data.begin(), po.data.end(), out);
}

int orc_richardson_lucy_u16(void* hp, const uint16_t* in, float* u, const float* psf, int sy, int sx, int sz, int niter, int use_ds,
                            int normalize, int resume, double tol, double* rel) {
    (void)tol;
    Handle* h = (Handle*)hp;
    PixelData<float> m(sy, sx, sz);
    std::copy(psf, psf + (size_t)sy * sx * sz, m.mesh.begin());
    ParticleData<uint16_t> pi;
    ParticleData<float> po;
    fill_parts(pi, in, h->apr.total_number_particles());
    if (resume) fill_parts(po, u, h->apr.total_number_particles());
    APRNumerics::richardson_lucy(h->apr, pi, po, m, niter, use_ds != 0, normalize != 0, resume != 0);
    std::copy(po.data.begin(), po.data.end(), u);
    if (rel) *rel = 0;
    return niter;
}

int orc_richardson_lucy_f32(void* hp, const float* in, float* u, const float* psf, int sy, int sx, int sz, int niter, int use_ds,
                            int normalize, int resume, double tol, double* rel) {
    (void)tol;
    Handle* h = (Handle*)hp;
    PixelData<float> m(sy, sx, sz);
    std::copy(psf, psf + (size_t)sy * sx * sz, m.mesh.begin());
    ParticleData<float> pi, po;
    fill_parts(pi, in, h->apr.total_number_particles());
    if (resume) fill_parts(po, u, h->apr.total_number_particles());
    APRNumerics::richardson_lucy(h->apr, pi, po, m, niter, use_ds != 0, normalize != 0, resume != 0);
    std::copy(po.data.begin(), po.data.end(), u);
    if (rel) *rel = 0;
    return niter;
}

/* APRNumerics::richardson_lucy_tv (APRNumerics.hpp:489-525, :549-572) */
}  // extern "C"
template <typename T>
static void ref_rl_tv(void* hp, const T* in, float* u, const float* psf, int sy, int sx, int sz, int niter, float reg, int use_ds,
                      int normalize, int resume) {
    Handle* h = (Handle*)hp;
    PixelData<float> m(sy, sx, sz);
    std::copy(psf, psf + (size_t)sy * sx * sz, m.mesh.begin());
    ParticleData<T> pi;
    ParticleData<float> po;
    fill_parts(pi, in, h->apr.total_number_particles());
    if (resume) fill_parts(po, u, h->apr.total_number_particles());
    APRNumerics::richardson_lucy_tv(h->apr, pi, po, m, niter, reg, use_ds != 0, normalize != 0, resume != 0);
    std::copy(po.data.begin(), po.data.end(), u);
}
extern "C" {
void orc_richardson_lucy_tv_u16(void* hp, const uint16_t* in, float* u, const float* psf, int sy, int sx, int sz, int niter, float reg,
                                int use_ds, int normalize, int resume) {
    ref_rl_tv(hp, in, u, psf, sy, sx, sz, niter, reg, use_ds, normalize, resume);
}
void orc_richardson_lucy_tv_f32(void* hp, const float* in, float* u, const float* psf, int sy, int sx, int sz, int niter, float reg,
                                int use_ds, int normalize, int resume) {
    ref_rl_tv(hp, in, u, psf, sy, sx, sz, niter, reg, use_ds, normalize, resume);
}

/* APRNumerics::gradient_cfd / gradient_sobel (APRNumerics.hpp:165-213); kind 0 cfd, 1 sobel */
}  // extern "C"
template <typename T>
static void ref_gradient(void* hp, const T* in, int kind, int dim, float delta, float* out) {
    Handle* h = (Handle*)hp;
    ParticleData<T> pi;
    ParticleData<float> po;
    fill_parts(pi, in, h->apr.total_number_particles());
    if (kind == 0) APRNumerics::gradient_cfd(h->apr, pi, po, dim, delta);
    else APRNumerics::gradient_sobel(h->apr, pi, po, dim, delta);
    std::copy(po.data.begin(), po.data.end(), out);
}
extern "C" {
void orc_gradient_u16(void* hp, const uint16_t* in, int kind, int dim, float delta, float* out) { ref_gradient(hp, in, kind, dim, delta, out); }
void orc_gradient_f32(void* hp, const float* in, int kind, int dim, float delta, float* out) { ref_gradient(hp, in, kind, dim, delta, out); }

/* APRNumerics::gradient_magnitude_cfd / _sobel (APRNumerics.hpp:216-283) */
}  // extern "C"
template <typename T>
static void ref_gradient_magnitude(void* hp, const T* in, int kind, const float* deltas, float* out) {
    Handle* h = (Handle*)hp;
    ParticleData<T> pi;
    ParticleData<float> po;
    fill_parts(pi, in, h->apr.total_number_particles());
    const std::vector<float> d = {deltas[0], deltas[1], deltas[2]};
    if (kind == 0) APRNumerics::gradient_magnitude_cfd(h->apr, pi, po, d);
    else APRNumerics::gradient_magnitude_sobel(h->apr, pi, po, d);
    std::copy(po.data.begin(), po.data.end(), out);
}
extern "C" {
void orc_gradient_magnitude_u16(void* hp, const uint16_t* in, int kind, const float* deltas, float* out) {
    ref_gradient_magnitude(hp, in, kind, deltas, out);
}
void orc_gradient_magnitude_f32(void* hp, const float* in, int kind, const float* deltas, float* out) {
    ref_gradient_magnitude(hp, in, kind, deltas, out);
}

/* APRNumerics::face_neighbour_filter (APRNumerics.hpp:305-355) and seperable_face_neighbour_filter
   (:286-302) on float particles; repeats < 0: the single-dimension filter along `dim` */
void orc_face_neighbour_filter_f32(void* hp, const float* in, const float* filter3, int dim, int repeats, float* out) {
    Handle* h = (Handle*)hp;
    ParticleData<float> pi, po;
    fill_parts(pi, in, h->apr.total_number_particles());
    const std::vector<float> f = {filter3[0], filter3[1], filter3[2]};
    if (repeats < 0) APRNumerics::face_neighbour_filter(h->apr, pi, po, f, dim);
    else APRNumerics::seperable_face_neighbour_filter(h->apr, pi, po, f, repeats);
    std::copy(po.data.begin(), po.data.end(), out);
}

void orc_gaussian_stencil(const float* sigma, const int* size, int normalize, float* out) {
    PixelData<float> m = APRStencil::create_gaussian_filter<float>({sigma[0], sigma[1], sigma[2]}, {size[0], size[1], size[2]},
                                                                   normalize != 0);
    std::copy(m.mesh.begin(), m.mesh.begin() + m.mesh.size(), out);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
