"""Error of every float result against a double-precision restatement (fp64_reference.py),
CPU only.

The GPU's float results are bit-equal to the oracle's (the -m gpu parity tests), so the
oracle's error against float64 is the GPU's.  Each check states two figures: the largest
absolute error over the largest magnitude of the float64 result, and the relative L2 error,
and asserts the 1e-5 relative bound the float stages are held to.  Run with
APRH_FP64_ERRORS=<path> to write the figures as JSON.

* APRConverter<float> stages on a 80x64x48 float volume, each from the float64 result of the
  stage before it (end to end from the input): smoothing spline, gradient magnitude with its
  2x2x2 maximum, the downsampled and inverse-B-spline intensity, the local intensity scale;
* fill_tree_mean of float particles;
* convolve_pencil, 5x5x5 Gaussian (its float64 ground truth per FilterTestHelpers.hpp:98-192);
* three Richardson-Lucy iterations (APRNumerics.hpp:549-593), every step in float64: tree
  means, both convolutions, the ratio and the product.
"""
import json
import os

import numpy as np
import pytest

import libapr_amd as A
import oracle_lib as O
import fp64_reference as R
from test_oracle_numerics import conv_gt, make_apr, particle_coords, stencil_stack, tree_coords

BOUND = 1e-5
FIGURES = {}


def _err(name, got, exp):
    got = np.asarray(got, np.float64).reshape(-1)
    exp = np.asarray(exp, np.float64).reshape(-1)
    d = np.abs(got - exp)
    e = {"max_abs_over_max": float(d.max() / np.abs(exp).max()),
         "rel_l2": float(np.linalg.norm(got - exp) / np.linalg.norm(exp))}
    FIGURES[name] = e
    print(f"fp64 {name}: max|err|/max|v| {e['max_abs_over_max']:.3e}  rel L2 {e['rel_l2']:.3e}")
    return e


@pytest.fixture(scope="module", autouse=True)
def _write_figures():
    yield
    path = os.environ.get("APRH_FP64_ERRORS")
    if path and FIGURES:
        with open(path, "w") as f:
            json.dump(FIGURES, f, indent=1, sort_keys=True)


def test_bspline_coefficients_match_the_float_setup():
    """the float64 coefficients for lambda 3, tol 1e-4: k0 = 18 and the float setup's values
    to float precision (rho 0.584374, b1 1.004104, b2 -0.341493, norm 0.113831)"""
    c = R.bspline_coeffs(3.0, 1e-4)
    assert c["k0"] == 18
    for k, v in (("rho", 0.584374), ("b1", 1.004104), ("b2", -0.341493), ("norm", 0.113831)):
        assert abs(c[k] - v) < 2e-6, (k, c[k])


def test_conversion_float_stages_against_fp64():
    shape = (48, 64, 80)                                   # z, x, y
    img = A.synth_blobs(shape, 5, 30).astype(np.float32)
    par = O.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)
    r = O.get_apr_f32(img, par, "oracle")
    s64 = R.bspline_smooth(img)
    g64 = R.gradient_ds(s64)
    lis64 = R.local_intensity_scale(R.inv_bspline(R.mean_ds(s64)))
    for name, got, exp in (("smoothed (float IIR, y x z)", r.intermediate(0), s64),
                           ("gradient magnitude, 2x2x2 max", r.intermediate(1), g64),
                           ("local intensity scale", r.intermediate(2), lis64)):
        e = _err(name, got, exp)
        assert e["max_abs_over_max"] < BOUND, name


@pytest.mark.parametrize("shape,nb,seed", [((32, 64, 64), 20, 0)])
def test_tree_mean_against_fp64(shape, nb, seed):
    img, o = make_apr(shape, nb, seed)
    t = o.tree()
    parts = (o.sample(img).astype(np.float32) * np.float32(1.37)).astype(np.float32)
    got = o.fill_tree_mean(parts)
    exp = R.tree_mean(o.dims, o.l_min, o.l_max, particle_coords(o), parts, tree_coords(o, t))
    e = _err("fill_tree_mean", got, exp)
    assert e["max_abs_over_max"] < BOUND


@pytest.mark.parametrize("shape,nb,seed", [((32, 64, 64), 20, 0)])
def test_convolution_and_rl_against_fp64(shape, nb, seed):
    img, o = make_apr(shape, nb, seed)
    t = o.tree()
    pc, tc = particle_coords(o), tree_coords(o, t)
    parts = o.sample(img).astype(np.float32)
    st = stencil_stack("gauss5", o, True)
    tree64 = R.tree_mean(o.dims, o.l_min, o.l_max, pc, parts, tc)
    e = _err("convolve_pencil 5x5x5 Gaussian", o.convolve_pencil(st, parts), conv_gt(o, st, parts, tree64, t, True))
    assert e["max_abs_over_max"] < BOUND
    # Richardson-Lucy: the 5x5x5 Gaussian as PSF, downsampled and normalised per level
    psf = O.gaussian_stencil((1.0, 1.0, 1.5), (5, 5, 5), True)
    st_f = O.downsampled_stencils(psf[::-1, ::-1, ::-1].copy(), o.l_max - o.l_min, True, "oracle")
    st_p = O.downsampled_stencils(psf, o.l_max - o.l_min, True, "oracle")
    k = 3
    got, _, _ = o.richardson_lucy(parts, psf, k, use_ds=True, normalize=True)
    u = np.ones(o.total)
    p64 = parts.astype(np.float64)
    for _ in range(k):
        blur = conv_gt(o, st_f, u, R.tree_mean(o.dims, o.l_min, o.l_max, pc, u, tc), t, True)
        rb = p64 / blur
        u = u * conv_gt(o, st_p, rb, R.tree_mean(o.dims, o.l_min, o.l_max, pc, rb, tc), t, True)
    e = _err("richardson_lucy, 3 iterations", got, u)
    assert e["max_abs_over_max"] < BOUND
