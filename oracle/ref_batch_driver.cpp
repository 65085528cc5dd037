/*
 * ref_batch_driver.cpp -- the reference's blocked conversion and blocked sampling
 * (APRConverterBatch<uint16_t>, APRConverterBatch.hpp:111-175; ParticleData::
 * sample_parts_from_img_blocked, ParticleData.hpp:140-141, 267-324), compiled
 * from the reference tree in place with the image's libtiff (both read their
 * input from a TIFF file, which TiffUtils::saveMeshAsTiff writes to a temporary
 * path first).  Built by oracle/Makefile into oracle/_ref/libref_batch.so only
 * where the reference tree and libtiff's header exist.  TEST INFRASTRUCTURE ONLY:
 * it pins the oracle's restatement of the blocked path.
 */
#include <cstdio>
#include <cstring>
#include <string>
#include <unistd.h>

/* the stage headers in the order APRConverter.hpp includes them (LocalParticleCellSet.hpp
   uses the pulling-scheme classes without including them) */
#include "algorithm/PullingScheme.hpp"
#include "algorithm/PullingSchemeSparse.hpp"
#include "algorithm/APRConverterBatch.hpp"
#include "data_structures/APR/particles/ParticleData.hpp"

#include "../include/aprh_params.h"

namespace {

struct Handle {
    APR apr;
    std::string tif;
};

std::string temp_tiff(const uint16_t* img, int64_t y, int64_t x, int64_t z) {
    char path[] = "/tmp/refbatchXXXXXX";
    const int fd = mkstemp(path);
    if (fd < 0) throw std::runtime_error("mkstemp failed");
    close(fd);
    std::string name = std::string(path) + ".tif";
    std::rename(path, name.c_str());
    PixelData<uint16_t> m;
    m.initWithValue((int)y, (int)x, (int)z, 0);
    std::copy(img, img + (size_t)(y * x * z), m.mesh.begin());
    TiffUtils::saveMeshAsTiff(name, m);
    return name;
}

}  // namespace

extern "C" {

void* refb_get_apr_by_block_u16(const uint16_t* img, int64_t y, int64_t x, int64_t z, const aprh_params* p, int64_t z_block_size,
                                int64_t ghost_z) {
    Handle* h = new Handle();
    h->tif = temp_tiff(img, y, x, z);
    APRConverterBatch<uint16_t> conv;
    conv.par.dx = p->dx; conv.par.dy = p->dy; conv.par.dz = p->dz;
    conv.par.psfx = p->psfx; conv.par.psfy = p->psfy; conv.par.psfz = p->psfz;
    conv.par.Ip_th = p->Ip_th; conv.par.lambda = p->lambda; conv.par.rel_error = p->rel_error;
    conv.par.sigma_th = p->sigma_th; conv.par.sigma_th_max = p->sigma_th_max; conv.par.grad_th = p->grad_th;
    conv.par.reflect_bc_lis = p->reflect_bc_lis != 0;
    conv.par.neighborhood_optimization = p->neighborhood_optimization != 0;
    conv.par.constant_intensity_scale = p->constant_intensity_scale != 0;
    conv.par.auto_parameters = false;
    conv.par.output_steps = false;
    conv.par.input_dir = "";
    conv.par.input_image_name = h->tif;
    conv.z_block_size = (int)z_block_size;
    conv.ghost_z = (int)ghost_z;
    conv.set_generate_linear(true);
    conv.set_sparse_pulling_scheme(false);
    if (!conv.get_apr(h->apr)) {
        std::remove(h->tif.c_str());
        delete h;
        return nullptr;
    }
    return h;
}

/* total, n_xz, n_lxz */
void refb_info(void* hp, int64_t* out) {
    Handle* h = (Handle*)hp;
    out[0] = (int64_t)h->apr.total_number_particles();
    out[1] = (int64_t)h->apr.linearAccess.xz_end_vec.size();
    out[2] = (int64_t)h->apr.linearAccess.level_xz_vec.size();
}

void refb_get(void* hp, uint64_t* level_xz, uint64_t* xz_end, uint16_t* y_vec) {
    Handle* h = (Handle*)hp;
    auto& L = h->apr.linearAccess;
    std::copy(L.level_xz_vec.begin(), L.level_xz_vec.end(), level_xz);
    std::copy(L.xz_end_vec.begin(), L.xz_end_vec.end(), xz_end);
    std::copy(L.y_vec.begin(), L.y_vec.end(), y_vec);
}

void refb_sample_blocked(void* hp, int64_t z_block_size, int64_t ghost_z, uint16_t* parts) {
    Handle* h = (Handle*)hp;
    ParticleData<uint16_t> pd;
    pd.sample_parts_from_img_blocked(h->apr, h->tif, (int)z_block_size, (int)ghost_z);
    std::copy(pd.data.begin(), pd.data.end(), parts);
}

void refb_free(void* hp) {
    Handle* h = (Handle*)hp;
    std::remove(h->tif.c_str());
    delete h;
}

}  // extern "C"
