/*
 * aprh_params.h -- plain-C parameter block for the APR conversion path.
 *
 * Field-for-field mirror of the reference's APRParameters
 * (src/algorithm/APRParameters.hpp:15-39).  Only the fields that influence
 * get_apr's numerical result are carried; the string fields (names,
 * directories, mask file) and output_steps have no effect on the result and
 * are handled by the host facade.
 */
#ifndef APRH_PARAMS_H
#define APRH_PARAMS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aprh_params {
    /* pixel spacing (APRParameters.hpp:16-18) */
    float dx, dy, dz;
    /* local-intensity-scale window (APRParameters.hpp:21-23) */
    float psfx, psfy, psfz;
    /* main pipeline parameters (APRParameters.hpp:26-31) */
    float Ip_th;
    float lambda;
    float rel_error;
    float sigma_th;
    float sigma_th_max;
    float grad_th;
    /* flags (APRParameters.hpp:35-39) */
    int reflect_bc_lis;             /* default 1 */
    int neighborhood_optimization;  /* default 1 */
    int constant_intensity_scale;   /* default 0 */
} aprh_params;

/* Defaults of the reference's APRParameters (APRParameters.hpp:15-39). */
static inline aprh_params aprh_params_default(void) {
    aprh_params p;
    p.dx = 1.0f; p.dy = 1.0f; p.dz = 1.0f;
    p.psfx = 2.0f; p.psfy = 2.0f; p.psfz = 2.0f;
    p.Ip_th = 0.0f; p.lambda = 3.0f; p.rel_error = 0.1f;
    p.sigma_th = 0.0f; p.sigma_th_max = 0.0f; p.grad_th = 0.0f;
    p.reflect_bc_lis = 1; p.neighborhood_optimization = 1; p.constant_intensity_scale = 0;
    return p;
}

#ifdef __cplusplus
}
#endif

#endif /* APRH_PARAMS_H */
