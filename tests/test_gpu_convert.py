"""GPU parity of the conversion path (get_apr + sample_image) against the CPU
oracle, through the C-ABI.  Bit-exact for every integer / index result and for
the float stage outputs (the kernels keep the reference's FP32 order)."""
import numpy as np
import pytest

import libapr_amd as A

pytestmark = pytest.mark.gpu

CASES = [
    # (z, x, y), n_blobs, seed
    ((32, 64, 64), 20, 0),
    ((61, 63, 65), 30, 1),
    ((3, 37, 200), 10, 2),
    ((128, 256, 256), 500, 0),   # the 256x256x128 conversion workload
    # long y rows: the top particle-cell level has rows of 512+ cells, which take the 64-lane
    # row fill and the multi-chunk row count used at 2048x2048x512
    ((6, 10, 1030), 4, 6),
    ((4, 8, 2048), 4, 7),
]


def _par(O):
    return O.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)


def _apar():
    return A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)


@pytest.mark.parametrize("shape,nb,seed", CASES)
def test_get_apr_matches_oracle(oracle, ctx, shape, nb, seed):
    img = A.synth_blobs(shape, seed, nb)
    ref = oracle.get_apr(img, _par(oracle), "oracle")
    conv = A.APRConverter(ctx)
    conv.par = _apar()
    apr = conv.get_apr(img)
    print(f"case {shape} P={apr.total} oracle P={ref.total}", flush=True)
    names = ["smoothed", "grad_raw", "lis_raw", "grad", "lis", "levels"]
    bad = []
    for w, nm in enumerate(names):
        g, o = apr.intermediate(w), ref.intermediate(w)
        nmis = int((g != o).sum())
        if nmis:
            idx = np.argwhere(g != o)[:3]
            bad.append(f"{nm}: {nmis}/{g.size} e.g. {idx.tolist()} gpu={g[tuple(idx[0])]} oracle={o[tuple(idx[0])]}")
    assert not bad, "\n".join(bad)
    assert apr.total == ref.total
    lxz, xze, yv = apr.structure()
    assert np.array_equal(lxz, ref.level_xz_vec)
    assert np.array_equal(xze, ref.xz_end_vec)
    assert np.array_equal(yv, ref.y_vec)
    parts = A.ParticleData().sample_image(apr, img)
    assert np.array_equal(parts, ref.sample(img))


@pytest.mark.parametrize("shape,nb,seed", CASES + [((4, 6, 5), 1, 3), ((17, 45, 128), 12, 4), ((9, 70, 192), 9, 5)])
def test_get_apr_and_sample_matches_separate_calls(oracle, ctx, shape, nb, seed):
    # one call that converts and samples (the row fill gathers the intensities) against
    # get_apr + sample_image and the oracle; host and device images; a buffer one particle too
    # small leaves the sampling to the caller
    img = A.synth_blobs(shape, seed, nb)
    conv = A.APRConverter(ctx)
    conv.par = _apar()
    ref_apr = conv.get_apr(img)
    ref_parts = A.ParticleData().sample_image(ref_apr, img)
    P = ref_apr.total
    print(f"case {shape} P={P} l_max={ref_apr.l_max}", flush=True)
    d_img = A.DeviceArray(img.nbytes, ctx)
    d_img.from_host(img)
    for src, dims in ((img, None), (d_img.ptr, shape[::-1])):
        parts = A.DeviceArray(P * 2 + 2, ctx)
        apr, sampled = conv.get_apr_and_sample(src, parts.ptr, P, dims=dims)
        assert sampled and apr.total == P
        lxz, xze, yv = apr.structure()
        rl, rx, ry = ref_apr.structure()
        assert np.array_equal(xze, rx) and np.array_equal(yv, ry) and np.array_equal(lxz, rl)
        assert np.array_equal(parts.to_host(np.uint16, P), ref_parts)
        apr.close()
        if P > 0:
            apr, sampled = conv.get_apr_and_sample(src, parts.ptr, P - 1, dims=dims)
            assert not sampled and apr.total == P
            apr.close()
        parts.free()
    assert np.array_equal(ref_parts, oracle.get_apr(img, _par(oracle), "oracle").sample(img))
    d_img.free()


@pytest.mark.parametrize("lo,hi,overflow", [(20000, 47000, False), (30000, 35500, False), (5, 40000, False),
                                             (2000, 64000, True)])
def test_get_apr_offset_not_guessed(oracle, ctx, lo, hi, overflow):
    # the y pass starts from the offset that images with max + min <= 65335 get (100) and runs
    # again with the true offset when the (min, max) it gathers say otherwise: the first two
    # images (offsets -100 and 17) take the second run, the third the first; the last overflows
    # the smoothing under either offset.  lambda 0.2 keeps the wide ranges below overflow.
    z, x, y = 24, 40, 128
    img = A.synth_blobs((z, x, y), 5, 40).astype(np.float64)
    img = (lo + (img - img.min()) * (hi - lo) / max(img.max() - img.min(), 1.0)).round().astype(np.uint16)
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10, lambda_=0.2)
    if overflow:
        with pytest.raises(A.AprhOverflowError):
            conv.get_apr(img)
        return
    ref = oracle.get_apr(img, oracle.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10, lambda_=0.2), "oracle")
    apr = conv.get_apr(img)
    for w in range(6):
        assert np.array_equal(apr.intermediate(w), ref.intermediate(w)), w
    assert apr.total == ref.total
    lxz, xze, yv = apr.structure()
    assert np.array_equal(xze, ref.xz_end_vec) and np.array_equal(yv, ref.y_vec)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
def test_bspline_axis(oracle, ctx, axis, dtype):
    rng = np.random.default_rng(axis)
    img = (rng.random((9, 23, 47)) * 2000 + 200).astype(dtype)
    g = img.copy()
    A.bspline_filt_rec(g, axis, 3.0, 1e-4, ctx)
    o = img.copy()
    lib = oracle.load("oracle")
    z, x, y = img.shape
    (lib.orc_bspline_u16 if dtype == np.uint16 else lib.orc_bspline_f32)(o.reshape(-1), y, x, z, axis, 3.0, 1e-4)
    assert np.array_equal(g.view(np.uint8), o.view(np.uint8))


@pytest.mark.parametrize("n", [64, 65, 66, 67, 68, 69, 75, 83, 84, 85, 100, 260])
@pytest.mark.parametrize("axis", [1, 2])
@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
def test_bspline_outer_chunked(oracle, ctx, n, axis, dtype):
    # x / z lines of 64 samples or more take the pass that saves causal states per chunk of
    # 16 steps instead of storing the causal output: every remainder of (n - 4) mod 16
    rng = np.random.default_rng(n + 7 * axis)
    shape = (n, 5, 24) if axis == 2 else (3, n, 24)
    img = (rng.random(shape) * 3000 + 100).astype(dtype)
    g = img.copy()
    A.bspline_filt_rec(g, axis, 3.0, 1e-4, ctx)
    o = img.copy()
    lib = oracle.load("oracle")
    z, x, y = img.shape
    (lib.orc_bspline_u16 if dtype == np.uint16 else lib.orc_bspline_f32)(o.reshape(-1), y, x, z, axis, 3.0, 1e-4)
    assert np.array_equal(g.view(np.uint8), o.view(np.uint8))


@pytest.mark.parametrize("shape", [(19, 21, 23), (17, 30, 26)])   # odd y: one sample per lane; even y: two
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("off", [1, 2, 3, 6])
def test_sat_mean_axis(oracle, ctx, axis, off, shape):
    rng = np.random.default_rng(7 + off)
    img = (rng.random(shape) * 100).astype(np.float32)
    g = img.copy()
    A.sat_mean(g, axis, off, ctx)
    o = img.copy()
    z, x, y = img.shape
    oracle.load("oracle").orc_sat_mean_f32(o.reshape(-1), y, x, z, axis, off)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))


@pytest.mark.parametrize("pattern", ["random", "checker", "ramps"])
def test_gradient_sqrt_exact(oracle, ctx, pattern):
    # extreme differences exercise the integer square root against the reference's double
    # sqrt + truncation; the checkerboard of 0 / 65535 reaches the largest sum of squares
    rng = np.random.default_rng(3)
    shape = (8, 16, 40)
    if pattern == "random":
        img = rng.integers(0, 65535, size=shape, dtype=np.uint16)
    elif pattern == "checker":
        zz, xx, yy = np.indices(shape)
        img = (((zz + xx + yy) % 2) * 65535).astype(np.uint16)
    else:
        zz, xx, yy = np.indices(shape)
        img = ((zz * 8191 + xx * 4093 + yy * 1637 + rng.integers(0, 3, size=shape)) % 65536).astype(np.uint16)
    g, m = A.gradient_ds(img, ctx=ctx)
    z, x, y = img.shape
    go = np.zeros_like(g)
    mo = np.zeros_like(m)
    lib = oracle.load("oracle")
    lib.orc_grad_mag_ds_u16(img.reshape(-1), go.reshape(-1), y, x, z, 1.0, 1.0, 1.0)
    lib.orc_downsample_mean_u16_f32(img.reshape(-1), mo.reshape(-1), y, x, z)
    assert np.array_equal(g, go)
    assert np.array_equal(m.view(np.uint32), mo.view(np.uint32))


F32_CASES = [((32, 64, 64), 20, 0, 0.731), ((29, 45, 61), 12, 3, 1.913), ((3, 37, 200), 10, 2, 0.05), ((64, 128, 192), 80, 6, 3.0)]


@pytest.mark.parametrize("shape,nb,seed,scale", F32_CASES)
def test_get_apr_f32_matches_oracle(oracle, ctx, shape, nb, seed, scale):
    """APRConverter<float>: every stage output, the structure and float sampling bit-equal"""
    img = (A.synth_blobs(shape, seed, nb).astype(np.float32) * np.float32(scale)).astype(np.float32)
    kw = dict(rel_error=0.1, sigma_th=50 * scale, Ip_th=120 * scale, grad_th=10 * scale)
    ref = oracle.get_apr_f32(img, oracle.params(**kw), "oracle")
    conv = A.APRConverter(ctx, image_type=np.float32)
    conv.par = A.APRParameters(**kw)
    apr = conv.get_apr(img)
    print(f"f32 case {shape} P={apr.total} oracle P={ref.total}", flush=True)
    assert apr.bspline_offset == 0.0 and apr.image_dtype == np.float32
    bad = []
    for w in range(6):
        g, o = apr.intermediate(w), ref.intermediate(w)
        assert g.dtype == o.dtype
        nmis = int((g.view(np.uint8) != o.view(np.uint8)).sum())
        if nmis:
            bad.append(f"intermediate {w}: {nmis} bytes differ")
    assert not bad, "\n".join(bad)
    lxz, xze, yv = apr.structure()
    assert apr.total == ref.total
    assert np.array_equal(xze, ref.xz_end_vec) and np.array_equal(yv, ref.y_vec)
    parts = A.ParticleData().sample_image(apr, img)
    assert np.array_equal(parts.view(np.uint32), ref.sample_f32(img).view(np.uint32))
