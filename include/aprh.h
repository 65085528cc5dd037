/*
 * aprh.h -- C-ABI of libaprh, the MI355X (gfx950) implementation of LibAPR's
 * APR hot path.  Plain pointers and sizes only; every entry point returns an
 * aprh_status (0 = success) and aprh_last_error() gives the message of the
 * last failure on the calling thread.
 *
 * Pointers: image / particle buffers may be host or device memory (the
 * library inspects them); device pointers must belong to the context's
 * device.  Arrays are y-fastest: idx = z*x_num*y_num + x*y_num + y
 * (reference src/data_structures/Mesh/PixelData.hpp:509,521).
 *
 * Every call is enqueued on the context's stream and is complete when it
 * returns (calls on one context are serialised).
 */
#ifndef APRH_H
#define APRH_H

#include <stddef.h>
#include <stdint.h>

#include "aprh_params.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum aprh_status {
    APRH_OK = 0,
    APRH_ERR_INVALID_ARGUMENT = 1, /* bad pointer / size / parameter                          */
    APRH_ERR_OVERFLOW = 2,         /* u16 IIR under/overflow: std::invalid_argument in the
                                      reference (ComputeGradient.hpp:409-412,595-598,775-778) */
    APRH_ERR_BAD_DIMS = 3,         /* get_apr returning false (APRConverter.hpp:86-91)        */
    APRH_ERR_UNSUPPORTED = 4,      /* std::runtime_error in the reference                     */
    APRH_ERR_OUT_OF_MEMORY = 5,
    APRH_ERR_HIP = 6               /* any other HIP runtime failure                           */
} aprh_status;

typedef struct aprh_context aprh_context;
typedef struct aprh_apr aprh_apr;

/* GenInfo summary of an APR (reference GenInfo.hpp:20-31) */
typedef struct aprh_apr_info {
    int32_t l_min, l_max;
    int64_t org_dims[3];            /* y, x, z                                       */
    uint64_t total_number_particles;
    uint64_t xz_end_size;           /* elements of xz_end_vec                        */
    uint64_t level_xz_size;         /* elements of level_xz_vec (= l_max + 2)        */
    float bspline_offset;
} aprh_apr_info;

const char* aprh_last_error(void);
int aprh_version(void);

/* ---- context (one HIP stream on one device) ---- */
int aprh_context_create(int device, aprh_context** out);
int aprh_context_destroy(aprh_context* ctx);
/* keep stage outputs (smoothed image, gradient, local scale, levels, particle-cell tree)
   inside each aprh_apr for inspection; replaces APRParameters::output_steps
   (APRConverter.hpp:231-244) */
int aprh_context_set_keep_intermediates(aprh_context* ctx, int keep);
/* per-stage device time (ms, HIP events on the context stream) of the last get_apr /
   sample_image; replaces APRConverter's timers (APRConverter.hpp:61-65).  Stages are
   named by aprh_stage_name(i), i < aprh_stage_count(). */
int aprh_context_set_timing(aprh_context* ctx, int on);
int aprh_context_stage_times(aprh_context* ctx, float* ms, int n);
int aprh_stage_count(void);
const char* aprh_stage_name(int i);
int aprh_context_release_cache(aprh_context* ctx);

/* ---- conversion ---- */
/* APRConverter<uint16_t>::get_apr (src/algorithm/APRConverter.hpp:591-600, get_apr_cpu :534-582) */
int aprh_get_apr_u16(aprh_context* ctx, const uint16_t* image, int64_t y_num, int64_t x_num, int64_t z_num,
                     const aprh_params* par, aprh_apr** out);
/* APRConverter::get_apr_custom_grad_scale (APRConverter.hpp:169-191): structure from a given
   downsampled gradient (u16) and local intensity scale (f32); intensity (f32, downsampled) is
   the Ip_th reference that get_apr would take from local_scale_temp2 */
int aprh_apr_from_grad_lis(aprh_context* ctx, const uint16_t* grad_ds, const float* lis_ds, const float* intensity_ds,
                           int64_t y_num, int64_t x_num, int64_t z_num, const aprh_params* par, float bspline_offset,
                           aprh_apr** out);
int aprh_apr_info_get(const aprh_apr* apr, aprh_apr_info* info);
/* copies LinearAccess (LinearAccess.hpp:33-35) out; any pointer may be NULL to skip */
int aprh_apr_copy_structure(const aprh_apr* apr, uint64_t* level_xz_vec, uint64_t* xz_end_vec, uint16_t* y_vec);
/* device pointers of the structure (valid until aprh_apr_destroy) */
int aprh_apr_device_structure(const aprh_apr* apr, const uint64_t** xz_end_vec, const uint16_t** y_vec);
/* kept stage outputs, see aprh_context_set_keep_intermediates.  which:
   0 smoothed u16[N], 1 gradient_step u16[Nd], 2 local_intensity_scale_step f32[Nd],
   3 thresholded gradient u16[Nd], 4 rescaled scale f32[Nd], 5 levels f32[Nd]  (Nd = ds volume) */
int aprh_apr_intermediate(const aprh_apr* apr, int which, void* dst);
/* final particle-cell tree of pulling-scheme level l (PullingScheme.hpp:83), u8 */
int aprh_apr_particle_cell_tree(const aprh_apr* apr, int level, uint8_t* dst, int64_t* dims3);
int aprh_apr_destroy(aprh_apr* apr);

/* ParticleData<uint16_t>::sample_image (ParticleData.hpp:122-127, :221-261) */
int aprh_sample_image_u16(aprh_context* ctx, const aprh_apr* apr, const uint16_t* image, uint16_t* parts);

/* ---- stage-level entry points (public methods of the reference's stage classes) ---- */
/* ComputeGradient::bspline_filt_rec_y/_x/_z (ComputeGradient.hpp:248-413, 415-599, 601-779); axis 0=y 1=x 2=z.
   In place.  Returns APRH_ERR_OVERFLOW when a rounded u16 value leaves [0,65535]. */
int aprh_bspline_filt_rec_u16(aprh_context* ctx, uint16_t* image, int64_t y_num, int64_t x_num, int64_t z_num, int axis,
                              float lambda, float tol);
int aprh_bspline_filt_rec_f32(aprh_context* ctx, float* image, int64_t y_num, int64_t x_num, int64_t z_num, int axis,
                              float lambda, float tol);
/* ComputeGradient::calc_inv_bspline_y/_x/_z (ComputeGradient.hpp:784-964), in place */
int aprh_inv_bspline_f32(aprh_context* ctx, float* image, int64_t y_num, int64_t x_num, int64_t z_num, int axis);
/* ComputeGradient::calc_bspline_fd_ds_mag (ComputeGradient.hpp:975-1051) fused with the
   gradient stage's mean downsample (ComputeGradient.hpp:95-98); mean_ds may be NULL.
   grad_ds must be zero-initialised like grad_temp (APRConverter.hpp:145-146). */
int aprh_gradient_ds_u16(aprh_context* ctx, const uint16_t* image, uint16_t* grad_ds, float* mean_ds, int64_t y_num,
                         int64_t x_num, int64_t z_num, float hx, float hy, float hz);
/* LocalIntensityScale::calc_sat_mean_y/_x/_z (LocalIntensityScale.hpp:311-494), in place */
int aprh_sat_mean_f32(aprh_context* ctx, float* image, int64_t y_num, int64_t x_num, int64_t z_num, int axis, int offset);
/* LocalIntensityScale::get_local_intensity_scale (LocalIntensityScale.hpp:26-150):
   intensity_ds -> scale_ds */
int aprh_local_intensity_scale(aprh_context* ctx, const float* intensity_ds, float* scale_ds, int64_t y_num, int64_t x_num,
                               int64_t z_num, const aprh_params* par);

/* ---- synthetic volumes (benchmark / test inputs; no reference counterpart) ---- */
/* background + blobs A*exp(-r^2/(2 sigma^2)) (r <= 4 sigma), optional shot noise; identical
   bytes on host and device */
int aprh_synth_blobs_host(uint16_t* out, int64_t y_num, int64_t x_num, int64_t z_num, uint64_t seed, int64_t n_blobs,
                          float sigma_min, float sigma_max, float amp_min, float amp_max, float background, int noise);
int aprh_synth_blobs_device(aprh_context* ctx, uint16_t* out, int64_t y_num, int64_t x_num, int64_t z_num, uint64_t seed,
                            int64_t n_blobs, float sigma_min, float sigma_max, float amp_min, float amp_max,
                            float background, int noise);

/* ---- device memory helpers for callers without their own allocator ---- */
int aprh_malloc(aprh_context* ctx, size_t bytes, void** out);
int aprh_free(aprh_context* ctx, void* p);
int aprh_memcpy(aprh_context* ctx, void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* APRH_H */
