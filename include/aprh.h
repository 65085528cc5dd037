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
    int32_t image_dtype;            /* APRH_DTYPE_U16 or APRH_DTYPE_F32: the converter's ImageType */
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
/* how APR convolutions on this context sum a stencil:
   APRH_CONV_EXACT      every tap in the reference's (z, x, y) order (APRFilter.hpp:721-733):
                        results bit-equal to the reference;
   APRH_CONV_SEPARABLE  stencils that are outer products of 1-D factors (Gaussian PSFs and
                        their downsampled stacks) run as three 1-D passes over each level's
                        piecewise-constant image; others as EXACT.  Results within float
                        rounding of EXACT (no reference counterpart; faster for large stencils). */
enum { APRH_CONV_EXACT = 0, APRH_CONV_SEPARABLE = 1 };
int aprh_context_set_conv_mode(aprh_context* ctx, int mode);
int aprh_stage_count(void);
const char* aprh_stage_name(int i);
int aprh_context_release_cache(aprh_context* ctx);

/* ---- conversion ---- */
/* APRConverter<uint16_t>::get_apr (src/algorithm/APRConverter.hpp:591-600, get_apr_cpu :534-582) */
int aprh_get_apr_u16(aprh_context* ctx, const uint16_t* image, int64_t y_num, int64_t x_num, int64_t z_num,
                     const aprh_params* par, aprh_apr** out);
/* get_apr followed by ParticleData<uint16_t>::sample_image of the same image (the usual pair,
   APRConverter.hpp:591-600 then ParticleData.hpp:122-127), in one call: the row fill that
   writes y_vec_ also gathers the particle intensities.  parts is device memory with room for
   parts_capacity particles; *sampled = 1 when they were written (the APR's particle count is at
   most parts_capacity), else 0 and the caller samples with aprh_sample_image_u16.  The APR and
   the intensities equal those of aprh_get_apr_u16 + aprh_sample_image_u16. */
int aprh_get_apr_sample_u16(aprh_context* ctx, const uint16_t* image, int64_t y_num, int64_t x_num, int64_t z_num,
                            const aprh_params* par, aprh_apr** out, uint16_t* parts, uint64_t parts_capacity, int32_t* sampled);
/* APRConverterBatch<uint16_t>::get_apr (src/algorithm/APRConverterBatch.hpp:44-175): the image in
   z blocks of z_block_size planes (the last takes the remainder) with up to ghost_z ghost planes
   on each side; each block runs the pixel pipeline on its own (get_apr_method_patch :186-276,
   its own B-spline offset, the whole image's l_max) and seeds its non-ghost planes into one
   particle-cell tree (fill_patch, PullingScheme.hpp:175-244); one pulling scheme and linear
   structure follow.  ghost_z >= z_num gives get_apr's result.  Rows a block would place past a
   level's z extent are skipped (the reference writes them beyond the level's rows). */
int aprh_get_apr_by_block_u16(aprh_context* ctx, const uint16_t* image, int64_t y_num, int64_t x_num, int64_t z_num,
                              const aprh_params* par, int64_t z_block_size, int64_t ghost_z, aprh_apr** out);
/* APRConverter<float>::get_apr (APRConverter.hpp:534-600 with ImageType = float): no B-spline
   offset (:218-220), float smoothing and gradient; same structure semantics */
int aprh_get_apr_f32(aprh_context* ctx, const float* image, int64_t y_num, int64_t x_num, int64_t z_num, const aprh_params* par,
                     aprh_apr** out);
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
   3 thresholded gradient u16[Nd], 4 rescaled scale f32[Nd], 5 levels f32[Nd]  (Nd = ds volume);
   0, 1 and 3 are f32 for an APR from aprh_get_apr_f32 */
int aprh_apr_intermediate(const aprh_apr* apr, int which, void* dst);
/* final particle-cell tree of pulling-scheme level l (PullingScheme.hpp:83), u8 */
int aprh_apr_particle_cell_tree(const aprh_apr* apr, int level, uint8_t* dst, int64_t* dims3);
int aprh_apr_destroy(aprh_apr* apr);

/* ---- batched conversion of many volumes of uneven sizes ----
   APRConverter<uint16_t>::get_apr (APRConverter.hpp:534-600) over a fixed list of volumes,
   recorded once into one HIP graph.  The plan owns each volume's u16 input buffer (write it
   through the pointer from aprh_batch_input) and its output APR; aprh_batch_run replays the
   graph and fills in the particle counts.  The APRs from aprh_batch_get are borrowed: valid
   until the next aprh_batch_run or aprh_batch_destroy, never passed to aprh_apr_destroy.
   Volumes need l_max >= 3 (a side above 4 pixels).  A batch has no reference counterpart:
   the reference converts one image per call. */
typedef struct aprh_batch aprh_batch;
int aprh_batch_create(aprh_context* ctx, int32_t n_volumes, const int64_t* dims3 /* y, x, z per volume */,
                      const aprh_params* par, aprh_batch** out);
int aprh_batch_input(aprh_batch* batch, int32_t i, uint16_t** dev_image);
int aprh_batch_run(aprh_context* ctx, aprh_batch* batch);
int aprh_batch_get(aprh_batch* batch, int32_t i, aprh_apr** apr);
int aprh_batch_destroy(aprh_batch* batch);
/* the plan's merged stage launches (one per stage over all volumes that take part) and the
   grid the autotuner chose for each: one block per tile, or 1-16 blocks per CU walking the
   tiles persistently, whichever ran fastest at plan creation */
typedef struct aprh_batch_stage {
    const char* name;     /* stage body (static string) */
    int32_t volumes;      /* volumes in this launch     */
    int64_t tiles;        /* work tiles over them       */
    uint32_t grid;        /* blocks launched            */
    uint32_t threads;     /* threads per block          */
    float tuned_ms;       /* time at the chosen grid    */
    const char* shape;    /* chosen tile shape (static string), NULL for stages with one shape */
    int32_t n_shapes;     /* tile-shape alternatives the autotuner timed */
} aprh_batch_stage;
int aprh_batch_stage_count(const aprh_batch* batch);
int aprh_batch_stage_info(const aprh_batch* batch, int32_t i, aprh_batch_stage* info);

/* ParticleData<uint16_t>::sample_image (ParticleData.hpp:122-127, :221-261) */
int aprh_sample_image_u16(aprh_context* ctx, const aprh_apr* apr, const uint16_t* image, uint16_t* parts);
/* ParticleData<uint16_t>::sample_parts_from_img_blocked (ParticleData.hpp:140-141, 267-362): per z
   block (as aprh_get_apr_by_block_u16) the mean pyramid of the block with its ghost planes,
   sampled at the particles of the block's non-ghost planes of every level; particles no block
   reaches are 0 */
int aprh_sample_image_blocked_u16(aprh_context* ctx, const aprh_apr* apr, const uint16_t* image, int64_t z_block_size,
                                  int64_t ghost_z, uint16_t* parts);
/* ParticleData<float>::sample_image of a float image (same mean pyramid, float means) */
int aprh_sample_image_f32(aprh_context* ctx, const aprh_apr* apr, const float* image, float* parts);
/* aprh_sample_image_u16 that allocates nothing: no counterpart in the reference.  For callers
   that cannot allocate device memory (stream capture, serving loops, harnesses that forbid
   hipMalloc).  image and parts must be device memory; the mean pyramid lives in a fixed
   device array of aprh_noalloc_scratch_bytes() bytes per device (64 MiB: images up to about
   235M voxels), APRH_ERR_UNSUPPORTED beyond that.  Calls that overlap on one device race. */
int aprh_sample_image_u16_noalloc(aprh_context* ctx, const aprh_apr* apr, const uint16_t* image, uint16_t* parts);
size_t aprh_noalloc_scratch_bytes(void);

/* ---- particle operations ---- */
/* element type of particle buffers (ParticleData<uint16_t> / ParticleData<float>) */
typedef enum aprh_dtype { APRH_DTYPE_U16 = 0, APRH_DTYPE_F32 = 1 } aprh_dtype;

/* tree (interior cell) summary: GenInfo::init_tree (GenInfo.hpp:68-99) and the tree
   LinearAccess (APR::initialize_tree_dense, APR.hpp:633-689) */
typedef struct aprh_tree_info {
    int32_t l_min, l_max;
    uint64_t total_number_tree_particles;
    uint64_t xz_end_size;
    uint64_t level_xz_size;
} aprh_tree_info;

/* outcome of aprh_richardson_lucy */
typedef struct aprh_rl_stats {
    int32_t iterations;             /* iterations run                                       */
    double last_rel_update;         /* ||u_k+1 - u_k||_2 / ||u_k||_2 of the last iteration   */
} aprh_rl_stats;

/* APR from a stored structure, as APRFile::read_apr restores LinearAccess
   (LinearAccess.hpp:33-35): xz_end_vec[xz_end_size] for the GenInfo of (y, x, z) and
   y_vec[total].  Host or device pointers; the structure is validated. */
int aprh_apr_from_structure(aprh_context* ctx, int64_t y_num, int64_t x_num, int64_t z_num, const uint64_t* xz_end_vec,
                            uint64_t xz_end_size, const uint16_t* y_vec, uint64_t total, aprh_apr** out);
/* APR::initialize_tree_dense (APR.hpp:633-689) on first use; tree GenInfo + sizes */
int aprh_apr_tree_info(aprh_context* ctx, aprh_apr* apr, aprh_tree_info* info);
/* tree LinearAccess (LinearAccess.hpp:67-122) out; any pointer may be NULL to skip */
int aprh_apr_copy_tree(aprh_context* ctx, aprh_apr* apr, uint64_t* level_xz_vec, uint64_t* xz_end_vec, uint16_t* y_vec);
/* APRTreeNumerics::fill_tree_mean (APRTreeNumerics.hpp:63-202): parts[total] -> tree[tree total] */
int aprh_fill_tree_mean(aprh_context* ctx, aprh_apr* apr, const void* parts, int dtype, float* tree_out);
/* APRStencil::get_downsampled_stencils (APRStencil.hpp:81-97, downsample_stencil :12-78), host
   only.  Stencil layout: y fastest, then x, then z (PixelData mesh order); sizes odd.
   With out == NULL only *needed (total floats) is written; dims_out gets (y, x, z) per level. */
int aprh_downsampled_stencils(const float* stencil, int32_t sy, int32_t sx, int32_t sz, int32_t nlevels, int normalize, float* out,
                              int64_t out_capacity, int32_t* dims_out, int64_t* needed);
/* APRStencil::create_gaussian_filter (APRStencil.hpp:173-228), host only: sigma and size
   per axis in (y, x, z) order; out[size_y*size_x*size_z], y fastest */
int aprh_gaussian_stencil(const float* sigma3, const int32_t* size3, int normalize, float* out);
/* APRStencil::create_mean_filter (APRStencil.hpp:236-244), host only */
int aprh_mean_stencil(const int32_t* size3, float* out);
/* APRFilter::convolve_pencil, vector-of-stencils overload (APRFilter.hpp:89-91, :1094-1216):
   stencils[0] applies to level l_max, stencils[k] to level l_max - k, the last one to every
   coarser level; dims holds (y, x, z) per stencil, all odd. */
int aprh_convolve_pencil(aprh_context* ctx, aprh_apr* apr, const float* stencils, const int32_t* dims, int32_t n_stencils,
                         const void* parts_in, int dtype, float* parts_out, int reflect);
/* APRFilter::convolve (APRFilter.hpp:67-69, :971-1077): same result as convolve_pencil */
int aprh_convolve(aprh_context* ctx, aprh_apr* apr, const float* stencils, const int32_t* dims, int32_t n_stencils,
                  const void* parts_in, int dtype, float* parts_out, int reflect);
/* APRFilter::convolve_pencil / convolve with one stencil (APRFilter.hpp:110-120, :139-149):
   get_downsampled_stencils(stencil, use_stencil_downsample ? l_max - l_min : 1, normalize) first */
int aprh_convolve_stencil(aprh_context* ctx, aprh_apr* apr, const float* stencil, int32_t sy, int32_t sx, int32_t sz,
                          const void* parts_in, int dtype, float* parts_out, int reflect, int use_stencil_downsample, int normalize);
/* APRNumerics::richardson_lucy (APRNumerics.hpp:130-134, :549-593): u (f32[total]) is set to 1
   unless resume; at most max_iter iterations; with tol > 0 it stops after the first iteration
   whose relative update ||u_k+1 - u_k|| / ||u_k|| is below tol (no reference counterpart). */
int aprh_richardson_lucy(aprh_context* ctx, aprh_apr* apr, const void* parts_in, int dtype, float* u, const float* psf, int32_t sy,
                         int32_t sx, int32_t sz, int32_t max_iter, int use_stencil_downsample, int normalize, int resume, double tol,
                         aprh_rl_stats* stats);
/* APRNumerics::richardson_lucy_tv (APRNumerics.hpp:142-148, :489-525, :549-572): Richardson-Lucy
   with total-variation regularisation; each iteration's update is divided by
   1 - reg_factor * div(grad u / |grad u|) (div_norm_grad, :431-486, unit pixel sizes).  u
   (f32[total]) is set to 1 unless resume; exactly number_iterations iterations. */
int aprh_richardson_lucy_tv(aprh_context* ctx, aprh_apr* apr, const void* parts_in, int dtype, float* u, const float* psf, int32_t sy,
                            int32_t sx, int32_t sz, int32_t number_iterations, float reg_factor, int use_stencil_downsample, int normalize,
                            int resume);
/* APRReconstruction::reconstruct_constant (APRReconstruction.hpp:80-120; with level_delta < 0
   the full-volume patch form :297-394, tree means from fill_tree_mean).  image has the
   dimensions of level l_max + level_delta, (y, x, z) with y fastest. */
int aprh_reconstruct_constant(aprh_context* ctx, aprh_apr* apr, const void* parts, int parts_dtype, int level_delta, void* image,
                              int image_dtype);
/* APRReconstruction::reconstruct_level (APRReconstruction.hpp:131-171): each pixel of the full
   volume (y, x, z, y fastest) takes the level of the particle covering it; image_dtype U16 or
   F32 (the reference writes any pixel type) */
int aprh_reconstruct_level(aprh_context* ctx, aprh_apr* apr, void* image, int image_dtype);
/* APRReconstruction::reconstruct_smooth(apr, img, parts, scale_d) (APRReconstruction.hpp:270-295):
   the constant reconstruction smoothed along y, x, z by running-sum means whose half-width at a
   pixel is min(floor(2^(l_max - level) / scale_d[axis]), 20) (calc_sat_adaptive_y/x/z, :528-876);
   scale3 = NULL means the reference's default {2, 2, 2} */
int aprh_reconstruct_smooth(aprh_context* ctx, aprh_apr* apr, const void* parts, int parts_dtype, const float* scale3, void* image,
                            int image_dtype);
/* APRNumerics::gradient_cfd (APRNumerics.hpp:165-184) and gradient_sobel (:187-213): the
   particles convolved (convolve_pencil, reflected boundary) with the central difference
   [-1, 0, 1] / (2 delta) along `dimension` (0 y, 1 x, 2 z), or with the Sobel stencil (that
   difference times [1/4, 1/2, 1/4] smoothing along the other axes; the 2-D form for 2-D
   images, the central difference below), divided by 2^d at level delta d (rescaled, not
   downsampled, stencils).  out: float32[total]. */
int aprh_gradient_cfd(aprh_context* ctx, aprh_apr* apr, const void* parts, int dtype, int dimension, float delta, float* out);
int aprh_gradient_sobel(aprh_context* ctx, aprh_apr* apr, const void* parts, int dtype, int dimension, float delta, float* out);
/* gradient_magnitude_cfd / _sobel (APRNumerics.hpp:216-283): sqrtf(g_y^2 + g_x^2 + g_z^2) with
   deltas[3] the pixel sizes in (y, x, z); axes of size 1 are left out */
int aprh_gradient_magnitude_cfd(aprh_context* ctx, aprh_apr* apr, const void* parts, int dtype, const float* deltas, float* out);
int aprh_gradient_magnitude_sobel(aprh_context* ctx, aprh_apr* apr, const void* parts, int dtype, const float* deltas, float* out);

/* APRNumerics::face_neighbour_filter (APRNumerics.hpp:305-355): out = filter3[1] * p, plus
   filter3[2] times the mean of the neighbours across the +dimension face and filter3[0] across
   the -dimension face (the particle itself where a face has none); dimension 0 y, 1 x, 2 z */
int aprh_face_neighbour_filter(aprh_context* ctx, aprh_apr* apr, const void* parts, int dtype, const float* filter3, int dimension,
                               float* out);
/* APRNumerics::seperable_face_neighbour_filter (APRNumerics.hpp:286-302): repeats x (y, x, z) */
int aprh_seperable_face_neighbour_filter(aprh_context* ctx, aprh_apr* apr, const void* parts, int dtype, const float* filter3,
                                         int32_t repeats, float* out);
/* APRNumerics::local_std (APRNumerics.hpp:357-389): the box mean (size3 in (y, x, z), stencils
   downsampled and normalised) of the particles and of their squares (tree values: the squared
   tree means), then sqrtf(max(E[x^2] - E[x]^2, 0)).  out: float32[total]. */
int aprh_local_std(aprh_context* ctx, aprh_apr* apr, const void* parts, int dtype, const int32_t* size3, float* out);

/* ReconPatch (APRReconstruction.hpp:20-66): a box in pixels of the image at level
   l_max + level_delta (level_delta > 0 upsamples by 2^level_delta); begins below 0 become 0,
   ends below 0 the full extent ceil(dim * 2^level_delta), larger ends are cut to it */
typedef struct aprh_patch {
    int32_t x_begin, x_end, y_begin, y_end, z_begin, z_end;
    int32_t level_delta;
} aprh_patch;
/* ReconPatch::check_limits (APRReconstruction.hpp:38-64): clamps the bounds in place;
   APRH_ERR_INVALID_ARGUMENT when a range is empty afterwards */
int aprh_patch_check_limits(const aprh_apr* apr, aprh_patch* patch);
/* reconstruct_constant over a patch (APRReconstruction.hpp:186-193 -> :296-394; tree values
   from fill_tree_mean) and reconstruct_level over a patch (:397-490).  image holds the clamped
   patch, (y_end - y_begin, x_end - x_begin, z_end - z_begin) with y fastest */
int aprh_reconstruct_constant_patch(aprh_context* ctx, aprh_apr* apr, const void* parts, int parts_dtype, const aprh_patch* patch,
                                    void* image, int image_dtype);
int aprh_reconstruct_level_patch(aprh_context* ctx, aprh_apr* apr, const aprh_patch* patch, void* image, int image_dtype);

/* ---- learned Richardson-Lucy warm start (no reference counterpart) ----
   A 3-D CNN: four valid 3x3x3 convolutions 1 -> 16 -> 16 -> 16 -> 1 channels, ReLU after
   the first three, residual output y = x + conv4.  A tile of side S + 8 gives an output of
   side S.  Layers 1-3 run on bf16 MFMA with f32 accumulation, layer 4 in f32.  Parameters
   (f32, aprh_cnn_param_count() values): W1[27][16] b1[16] W2[27][16][16] b2[16]
   W3[27][16][16] b3[16] W4[27][16] b4, tap = (dz*3 + dx)*3 + dy. */
typedef struct aprh_cnn aprh_cnn;
int aprh_cnn_param_count(void);
/* He-initialised network (seeded); Adam state zero */
int aprh_cnn_create(aprh_context* ctx, uint64_t seed, aprh_cnn** out);
int aprh_cnn_destroy(aprh_cnn* net);
int aprh_cnn_get_params(aprh_context* ctx, aprh_cnn* net, float* out);
/* replaces the parameters and resets the Adam state */
int aprh_cnn_set_params(aprh_context* ctx, aprh_cnn* net, const float* in);
/* in[T][(S+8)^3] -> out[T][S^3] (f32, y fastest within a tile) */
int aprh_cnn_forward_tiles(aprh_context* ctx, aprh_cnn* net, const float* in, int32_t T, int32_t S, float* out);
/* one Adam step (lr 0: gradient only) on the mean squared error against target[T][S^3];
   loss (before the step) and, optionally, the gradient (param order) are returned */
int aprh_cnn_train_step(aprh_context* ctx, aprh_cnn* net, const float* in, const float* target, int32_t T, int32_t S, float lr,
                        float* loss, float* grads_out);
/* tiles of side size + 2*halo at origins (y, x, z per tile) minus halo, reflect-padded at the
   volume border, divided by scale */
int aprh_cnn_extract_tiles(aprh_context* ctx, const void* image, int dtype, int64_t y_num, int64_t x_num, int64_t z_num,
                           const int32_t* origins, int32_t T, int32_t size, int32_t halo, float scale, float* out);
/* the network over a whole image in 64^3 tiles (4-voxel reflect halo at the border);
   the image is divided by scale (<= 0: its mean) and the result multiplied back */
int aprh_cnn_apply(aprh_context* ctx, aprh_cnn* net, const void* image, int dtype, int64_t y_num, int64_t x_num, int64_t z_num,
                   float scale, float* out);
/* warm start for aprh_richardson_lucy(..., resume = 1): the network applied to the
   piecewise-constant reconstruction of the particles (normalised by its mean), floored at
   1e-3 of the mean, sampled back onto the particles (ParticleData::sample_image) */
int aprh_cnn_warm_start(aprh_context* ctx, aprh_cnn* net, aprh_apr* apr, const void* parts, int dtype, float* u0);
/* optional last step of aprh_cnn_warm_start: the estimate convolved (convolve_pencil, stencils
   downsampled per level and normalised, reflected boundary) with the Gaussian of
   APRStencil::create_gaussian_filter(sigma3, size3) in (y, x, z) order.  It removes the
   network's fine-scale error, which RL iterations do not correct, at the cost of a few more
   iterations.  Sizes odd in 1..15; NULL pointers switch it off (the default). */
int aprh_cnn_set_post_filter(aprh_cnn* net, const float* sigma3, const int32_t* size3);
/* mean of n u16 / f32 values (host or device), accumulated in double */
int aprh_mean(aprh_context* ctx, const void* data, int dtype, uint64_t n, double* mean);

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

/* deconvolution workload: gt = background + blobs (f32, device), blurred by a separable
   Gaussian PSF (sigma per axis y, x, z; psf_taps odd, normalised factors, reflective ends),
   then shot noise into noisy (u16, device) */
int aprh_synth_deconv_device(aprh_context* ctx, float* gt, uint16_t* noisy, int64_t y_num, int64_t x_num, int64_t z_num,
                             uint64_t seed, int64_t n_blobs, float sigma_min, float sigma_max, float amp_min, float amp_max,
                             float background, const float* psf_sigma3, int32_t psf_taps);
/* the same deconvolution workload computed on the host (gt and noisy host memory), byte for
   byte equal to aprh_synth_deconv_device */
int aprh_synth_deconv_host(float* gt, uint16_t* noisy, int64_t y_num, int64_t x_num, int64_t z_num, uint64_t seed,
                           int64_t n_blobs, float sigma_min, float sigma_max, float amp_min, float amp_max, float background,
                           const float* psf_sigma3, int32_t psf_taps);

/* ---- device memory helpers for callers without their own allocator ---- */
int aprh_malloc(aprh_context* ctx, size_t bytes, void** out);
int aprh_free(aprh_context* ctx, void* p);
int aprh_memcpy(aprh_context* ctx, void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* APRH_H */
