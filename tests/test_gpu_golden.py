"""The reference's known answers and pipeline fixtures, run through the HIP
library's C-ABI on the GPU (same checks as tests/test_oracle_golden.py)."""
import numpy as np
import pytest

import libapr_amd as A
import oracle_lib as O
from golden import kats
from test_oracle_golden import check_structure_against_images, fixture, mesh, zyx

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", range(len(kats.BSPLINE)))
def test_bspline_kat_gpu(ctx, case):
    dt, dims, fill, sets, axis, expect, eps = kats.BSPLINE[case]
    m = mesh(dims, fill, sets, np.uint16 if dt == "u16" else np.float32)
    A.bspline_filt_rec(m, axis, 3.0, 1e-4, ctx)
    np.testing.assert_allclose(zyx(m).astype(np.float64), np.array(expect, np.float64), atol=eps)


@pytest.mark.parametrize("dims,axis", kats.BSPLINE_CONSTANT)
@pytest.mark.parametrize("dt", ["u16", "f32"])
def test_bspline_constant_gpu(ctx, dims, axis, dt):
    m = mesh(dims, 100, [], np.uint16 if dt == "u16" else np.float32)
    A.bspline_filt_rec(m, axis, 3.0, 1e-4, ctx)
    np.testing.assert_allclose(m.astype(np.float64), 100.0, atol=0.01)


@pytest.mark.parametrize("case", range(len(kats.INV_BSPLINE)))
def test_inv_bspline_kat_gpu(ctx, case):
    dims, init, axis, expect = kats.INV_BSPLINE[case]
    y, x, z = dims
    m = np.transpose(np.array(init, np.float32).reshape(z, y, x), (0, 2, 1)).copy()
    A.inv_bspline(m, axis, ctx)
    np.testing.assert_allclose(zyx(m), expect, atol=0.01)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("case", [1, 2, 3])     # offsets 1 and 2 (the GPU means need offset >= 1)
def test_sat_mean_kat_gpu(ctx, axis, case):
    off, vin, expect = kats.SAT_MEAN[case]
    dims = [1, 1, 1]
    dims[axis] = 8
    y, x, z = dims
    m = np.array(vin, np.float32).reshape(z, x, y).copy()
    A.sat_mean(m, axis, off, ctx)
    np.testing.assert_allclose(m.reshape(-1), expect, atol=0.05)


@pytest.mark.parametrize("case", ["sphere_120", "sphere_diff_dims"])
def test_pipeline_fixture_gpu(ctx, case):
    f = fixture(case)
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(Ip_th=0.0, rel_error=float(f["attr_rel_error"]), lambda_=float(f["attr_lambda"]),
                               sigma_th_max=max(float(f["attr_sigma_th_max"]), 0.1),
                               sigma_th=max(float(f["attr_sigma_th"]), 0.1), grad_th=float(f["attr_grad_th"]))
    img = np.ascontiguousarray(f["original"])
    apr = conv.get_apr(img)
    g = apr.intermediate(1).astype(np.int64)
    s = apr.intermediate(2).astype(np.float64)
    assert int((np.abs(g - f["gradient_saved"].astype(np.int64)) > 1).sum()) == 0
    assert int((np.abs(s - f["scale_saved"].astype(np.float64)) > 1).sum()) == 0
    Z, X, Y = img.shape
    inten = np.ones_like(f["scale_saved"], dtype=np.float32)
    ap2 = conv.get_apr_custom_grad_scale(f["gradient_saved"], f["scale_saved"], inten, (Y, X, Z), 0.0)
    assert ap2.total == int(f["attr_total_number_particles"])
    lxz, xze, yv = ap2.structure()
    parts = A.ParticleData().sample_image(ap2, img)
    assert check_structure_against_images(f, ap2.l_min, ap2.l_max, xze, lxz, yv, parts) == 0


def test_overflow_detection_gpu(ctx):
    f = fixture("sphere_120")
    img = np.ascontiguousarray(f["original"]).copy()
    img.reshape(-1)[img.size // 2] = 65535 - 100
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(Ip_th=0.0, rel_error=0.1, lambda_=2.0, sigma_th=0.1, sigma_th_max=0.1, grad_th=10)
    with pytest.raises(A.AprhOverflowError):
        conv.get_apr(img)


@pytest.mark.parametrize("shape", [(8, 16, 128), (8, 24, 40), (5, 9, 13)])   # LDS, 16-byte and scalar y passes
def test_overflow_detection_spike(ctx, shape):
    """a lone spike near 65535 in a dark volume: the smoothing's negative side lobes leave [0, 65535]"""
    img = np.zeros(shape, np.uint16)
    z, x, y = shape
    img[z // 2, x // 2, y // 2] = 65435
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(lambda_=2.0)
    with pytest.raises(A.AprhOverflowError):
        conv.get_apr(img)


@pytest.mark.parametrize("shape,nb,seed", [((2, 3, 4), 1, 4), ((6, 6, 6), 2, 10), ((1, 1, 40), 2, 3), ((24, 1, 24), 4, 11)])
def test_tiny_volumes_gpu(oracle, ctx, shape, nb, seed):
    """tiny and degenerate shapes (l_max <= 2, constant local scale, 1-wide axes) against the oracle"""
    img = A.synth_blobs(shape, seed, nb)
    par = dict(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)
    ref = O.get_apr(img, O.params(**par), "oracle")
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(**par)
    apr = conv.get_apr(img)
    lxz, xze, yv = apr.structure()
    assert apr.total == ref.total
    assert np.array_equal(xze, ref.xz_end_vec) and np.array_equal(yv, ref.y_vec)
    assert np.array_equal(A.ParticleData().sample_image(apr, img), ref.sample(img))
