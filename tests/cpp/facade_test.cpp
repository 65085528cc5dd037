// C++ drop-in check: the apr_amd.hpp facade (reference-shaped API) against the
// CPU oracle (liboracle.so), on a small synthetic volume.  Exit status 0 and
// "facade ok" on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../include/apr_amd.hpp"

extern "C" {
void* orc_get_apr_u16(const uint16_t* img, int64_t y, int64_t x, int64_t z, const aprh_params* par, int keep, char* errbuf, int errlen);
void orc_apr_info(void* h, int64_t* out);
void orc_apr_get(void* h, uint64_t* level_xz, uint64_t* xz_end, uint16_t* y_vec);
void orc_sample_image_u16(void* h, const uint16_t* img, uint16_t* parts);
void orc_fill_tree_mean_u16(void* h, const uint16_t* parts, float* tree);
void orc_tree_info(void* h, int64_t* out);
int64_t orc_downsampled_stencils(const float* st, int sy, int sx, int sz, int nlevels, int normalize, float* out, int64_t cap,
                                 int* dims_out);
void orc_convolve_pencil_u16(void* h, const float* stencils, const int* dims, int n, const uint16_t* in, float* out, int reflect);
int orc_richardson_lucy_u16(void* h, const uint16_t* in, float* u, const float* psf, int sy, int sx, int sz, int niter, int use_ds,
                            int normalize, int resume, double tol, double* rel);
void orc_reconstruct(void* h, const void* parts, int pdt, int level_delta, void* img, int idt);
void orc_local_std_u16(void* h, const uint16_t* in, const int* size, float* out);
void orc_richardson_lucy_tv_u16(void* hp, const uint16_t* in, float* u, const float* psf, int sy, int sx, int sz, int niter, float reg,
                                int use_ds, int normalize, int resume);
void orc_reconstruct_smooth(void* hp, const void* parts, int pdt, const float* scale3, void* out, int idt);
void orc_face_neighbour_filter_f32(void* hp, const float* in, const float* filter3, int dim, int repeats, float* out);
void* orc_get_apr_by_block_u16(const uint16_t* img, int64_t y, int64_t x, int64_t z, const aprh_params* par, int64_t z_block_size,
                               int64_t ghost_z, char* errbuf, int errlen);
void orc_sample_image_blocked_u16(void* hp, const uint16_t* img, int64_t z_block_size, int64_t ghost_z, uint16_t* parts);
void orc_apr_free(void* h);
}

using namespace apr_amd;

static int failures = 0;
#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
            ++failures;                                                  \
        }                                                                \
    } while (0)

template <typename T>
static bool same_bytes(const std::vector<T>& a, const T* b, size_t n) {
    return a.size() == n && (n == 0 || std::memcmp(a.data(), b, n * sizeof(T)) == 0);
}

int main() {
    const int64_t y = 96, x = 72, z = 40;
    PixelData<uint16_t> img(y, x, z);
    if (aprh_synth_blobs_host(img.mesh.data(), y, x, z, 0, 24, 2.0f, 6.0f, 500.0f, 3000.0f, 100.0f, 1) != APRH_OK) return 2;

    Context ctx(0);
    APRConverter<uint16_t> conv(ctx);
    conv.par.rel_error = 0.1f;
    conv.par.sigma_th = 50.0f;
    conv.par.Ip_th = 120.0f;
    conv.par.grad_th = 10.0f;
    APR apr;
    CHECK(conv.get_apr(apr, img));

    char err[256];
    void* o = orc_get_apr_u16(img.mesh.data(), y, x, z, &conv.par, 0, err, sizeof(err));
    CHECK(o != nullptr);
    int64_t info[8];
    orc_apr_info(o, info);
    const uint64_t P = apr.total_number_particles();
    std::vector<uint64_t> lxz, xze;
    std::vector<uint16_t> yv;
    apr.linear_access(lxz, xze, yv);
    std::vector<uint64_t> olxz(lxz.size()), oxze(xze.size());
    std::vector<uint16_t> oyv(P);
    orc_apr_get(o, olxz.data(), oxze.data(), oyv.data());
    CHECK(same_bytes(xze, oxze.data(), oxze.size()));
    CHECK(same_bytes(yv, oyv.data(), oyv.size()));
    CHECK(same_bytes(lxz, olxz.data(), olxz.size()));

    ParticleData<uint16_t> parts;
    parts.sample_image(apr, img);
    std::vector<uint16_t> oparts(P);
    orc_sample_image_u16(o, img.mesh.data(), oparts.data());
    CHECK(same_bytes(parts.data, oparts.data(), P));

    // tree means
    ParticleData<float> tree;
    APRTreeNumerics::fill_tree_mean(apr, parts, tree);
    int64_t tinfo[5];
    orc_tree_info(o, tinfo);
    std::vector<float> otree((size_t)tinfo[2]);
    orc_fill_tree_mean_u16(o, oparts.data(), otree.data());
    CHECK(same_bytes(tree.data, otree.data(), otree.size()));

    // 5x5x5 Gaussian convolution, single-stencil overload (downsampled + normalised stack)
    PixelData<float> g5 = APRStencil::create_gaussian_filter({1.0f, 1.0f, 1.0f}, {5, 5, 5}, true);
    ParticleData<float> out;
    APRFilter::convolve_pencil(apr, g5, parts, out, true, true, true);
    const int nl = apr.level_max() - apr.level_min();
    const int64_t tot = orc_downsampled_stencils(g5.mesh.data(), 5, 5, 5, nl, 1, nullptr, 0, nullptr);
    std::vector<float> flat((size_t)tot);
    std::vector<int> dims(3 * (size_t)nl);
    orc_downsampled_stencils(g5.mesh.data(), 5, 5, 5, nl, 1, flat.data(), tot, dims.data());
    std::vector<float> oout(P);
    orc_convolve_pencil_u16(o, flat.data(), dims.data(), nl, oparts.data(), oout.data(), 1);
    CHECK(same_bytes(out.data, oout.data(), P));
    // the vector overload and APRFilter::convolve agree
    std::vector<PixelData<float>> stack;
    APRStencil::get_downsampled_stencils(g5, stack, nl, true);
    ParticleData<float> out2;
    APRFilter::convolve(apr, stack, parts, out2, true);
    CHECK(same_bytes(out2.data, oout.data(), P));

    // Richardson-Lucy, 2 iterations, 9x9x9 PSF
    PixelData<float> psf = APRStencil::create_gaussian_filter({1.0f, 1.0f, 2.0f}, {9, 9, 9}, true);
    ParticleData<float> u;
    aprh_rl_stats st = APRNumerics::richardson_lucy(apr, parts, u, psf, 2, true, true);
    std::vector<float> ou(P, 1.0f);
    double rel = 0;
    orc_richardson_lucy_u16(o, oparts.data(), ou.data(), psf.mesh.data(), 9, 9, 9, 2, 1, 1, 0, 0.0, &rel);
    CHECK(st.iterations == 2);
    CHECK(same_bytes(u.data, ou.data(), P));

    // reconstruction at full resolution and one level down
    for (int d = 0; d >= -1; --d) {
        PixelData<uint16_t> rec;
        APRReconstruction::reconstruct_constant(apr, rec, parts, d);
        std::vector<uint16_t> orec(rec.mesh.size());
        orc_reconstruct(o, oparts.data(), 0, d, orec.data(), 0);
        CHECK(same_bytes(rec.mesh, orec.data(), orec.size()));
    }

    // level image: the constant reconstruction of particles valued by their levels
    {
        std::vector<uint64_t> lxz, xze;
        std::vector<uint16_t> yv;
        apr.linear_access(lxz, xze, yv);
        std::vector<uint16_t> lev(P, 0);
        for (size_t l = 1; l + 1 < lxz.size(); ++l) {
            if (lxz[l + 1] == lxz[l]) continue;
            const uint64_t b = lxz[l] > 0 ? xze[lxz[l] - 1] : 0, e = xze[lxz[l + 1] - 1];
            for (uint64_t i = b; i < e; ++i) lev[i] = (uint16_t)l;
        }
        PixelData<uint16_t> rl;
        APRReconstruction::reconstruct_level(apr, rl);
        std::vector<uint16_t> orl((size_t)(y * x * z));
        orc_reconstruct(o, lev.data(), 0, 0, orl.data(), 0);
        CHECK(same_bytes(rl.mesh, orl.data(), orl.size()));
        PixelData<float> rlf;
        APRReconstruction::reconstruct_level(apr, rlf);
        bool eq = rlf.mesh.size() == orl.size();
        for (size_t i = 0; eq && i < orl.size(); ++i) eq = rlf.mesh[i] == (float)orl[i];
        CHECK(eq);

        // patches: a box of the full-resolution images (the end past the image is clamped)
        std::vector<uint16_t> orec((size_t)(y * x * z));
        orc_reconstruct(o, oparts.data(), 0, 0, orec.data(), 0);
        ReconPatch pp;
        pp.x_begin = 5, pp.x_end = 60, pp.y_begin = -2, pp.y_end = 500, pp.z_begin = 3, pp.z_end = 17;
        PixelData<uint16_t> pc, pl;
        APRReconstruction::reconstruct_constant(apr, pc, parts, pp);
        APRReconstruction::reconstruct_level(apr, pl, pp);
        CHECK(pp.y_begin == 0 && pp.y_end == y && pc.y_num == y && pc.x_num == 55 && pc.z_num == 14);
        bool peq = pc.mesh.size() == (size_t)(55 * 14 * y) && pl.mesh.size() == pc.mesh.size();
        for (int zz = 0; peq && zz < 14; ++zz)
            for (int xx = 0; peq && xx < 55; ++xx)
                for (int yy = 0; peq && yy < y; ++yy) {
                    const size_t src = ((size_t)(zz + 3) * x + (xx + 5)) * y + yy, dst = ((size_t)zz * 55 + xx) * y + yy;
                    peq = pc.mesh[dst] == orec[src] && pl.mesh[dst] == orl[src];
                }
        CHECK(peq);
        // an empty range: a message, the image untouched
        ReconPatch bad;
        bad.x_begin = 9, bad.x_end = 9;
        APRReconstruction::reconstruct_level(apr, pl, bad);
        CHECK(pl.mesh.size() == pc.mesh.size());
    }

    // local standard deviation in a 3x3x3 window
    {
        ParticleData<float> ls;
        APRNumerics::local_std(apr, parts, ls);
        std::vector<float> ols(P);
        const int sz3[3] = {3, 3, 3};
        orc_local_std_u16(o, oparts.data(), sz3, ols.data());
        CHECK(same_bytes(ls.data, ols.data(), P));
    }

    // total-variation Richardson-Lucy, 2 iterations
    {
        PixelData<float> psf = APRStencil::create_gaussian_filter({1.0f, 1.0f, 1.5f}, {5, 5, 5}, true);
        ParticleData<float> tv;
        APRNumerics::richardson_lucy_tv(apr, parts, tv, psf, 2, 0.01f, true, true, false);
        std::vector<float> otv(P, 1.0f);
        orc_richardson_lucy_tv_u16(o, oparts.data(), otv.data(), psf.mesh.data(), 5, 5, 5, 2, 0.01f, 1, 1, 0);
        CHECK(same_bytes(tv.data, otv.data(), P));
    }

    // smooth reconstruction, default scale {2, 2, 2}, float image
    {
        PixelData<float> rs;
        APRReconstruction::reconstruct_smooth(apr, rs, parts);
        std::vector<float> ors((size_t)(y * x * z));
        const float sc[3] = {2.0f, 2.0f, 2.0f};
        orc_reconstruct_smooth(o, oparts.data(), 0, sc, ors.data(), 1);
        CHECK(same_bytes(rs.mesh, ors.data(), ors.size()));
    }

    // face-neighbour filters on float particles
    {
        ParticleData<float> pf, ff;
        pf.init(P);
        for (size_t i = 0; i < P; ++i) pf.data[i] = (float)oparts[i];
        const std::vector<float> w = {0.25f, 0.5f, 0.25f};
        std::vector<float> off(P);
        APRNumerics::face_neighbour_filter(apr, pf, ff, w, 1);
        orc_face_neighbour_filter_f32(o, pf.data.data(), w.data(), 1, -1, off.data());
        CHECK(same_bytes(ff.data, off.data(), P));
        APRNumerics::seperable_face_neighbour_filter(apr, pf, ff, w, 2);
        orc_face_neighbour_filter_f32(o, pf.data.data(), w.data(), 0, 2, off.data());
        CHECK(same_bytes(ff.data, off.data(), P));
    }

    // blocked conversion and sampling: 16-plane blocks with 4 ghost planes
    {
        APRConverterBatch bconv(ctx);
        bconv.par = conv.par;
        bconv.z_block_size = 16;
        bconv.ghost_z = 4;
        APR bapr;
        CHECK(bconv.get_apr(bapr, img));
        void* ob = orc_get_apr_by_block_u16(img.mesh.data(), y, x, z, &conv.par, 16, 4, err, sizeof(err));
        CHECK(ob != nullptr);
        if (ob) {
            int64_t binfo[8];
            orc_apr_info(ob, binfo);
            const size_t bP = (size_t)binfo[2];
            CHECK(bapr.total_number_particles() == bP);
            std::vector<uint64_t> blx, bxz, olx((size_t)binfo[4]), oxz((size_t)binfo[3]);
            std::vector<uint16_t> byv, oyv(bP);
            bapr.linear_access(blx, bxz, byv);
            orc_apr_get(ob, olx.data(), oxz.data(), oyv.data());
            CHECK(blx == olx && bxz == oxz && byv == oyv);
            ParticleData<uint16_t> bp;
            bp.sample_parts_from_img_blocked(bapr, img, 16, 4);
            std::vector<uint16_t> obp(bP);
            orc_sample_image_blocked_u16(ob, img.mesh.data(), 16, 4, obp.data());
            CHECK(same_bytes(bp.data, obp.data(), bP));
            orc_apr_free(ob);
        }
    }

    // gradients: a constant field has none; a bad dimension throws std::invalid_argument
    {
        ParticleData<uint16_t> cst;
        cst.init(P);
        std::fill(cst.data.begin(), cst.data.end(), (uint16_t)777);
        ParticleData<float> gm;
        APRNumerics::gradient_magnitude_sobel(apr, cst, gm);
        bool zero = gm.data.size() == P;
        for (float v : gm.data) zero = zero && v == 0.0f;
        CHECK(zero);
        bool threw_dim = false;
        try {
            APRNumerics::gradient_cfd(apr, cst, gm, 3);
        } catch (std::invalid_argument&) {
            threw_dim = true;
        }
        CHECK(threw_dim);
    }

    // error conventions
    PixelData<uint16_t> empty;
    APR a2;
    CHECK(!conv.get_apr(a2, empty));
    bool threw = false;
    try {
        PixelData<float> even(4, 3, 3);
        APRFilter::convolve_pencil(apr, even, parts, out);
    } catch (std::runtime_error&) {
        threw = true;
    }
    CHECK(threw);

    orc_apr_free(o);
    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("facade ok: %lldx%lldx%lld -> %llu particles, bit-exact vs oracle\n", (long long)y, (long long)x, (long long)z,
                (unsigned long long)P);
    return 0;
}
