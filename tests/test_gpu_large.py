"""GPU parity at the benchmark workloads' full sizes.

The CPU oracle's results on these workloads are stored as digests (xxh3-64 of the
array bytes, plus values at fixed strides) in tests/golden/large_workloads.json,
written by tests/golden/make_large_fixtures.py from the host generators.  Each test
first checks that its device-generated input has the digest of the host-generated one,
then compares every result byte for byte:

  * 2048x2048x512 u16 (N = 2.15e9 voxels > 2^31, P ~ 1.7e9): get_apr + sample_image in
    one call (level_xz_vec, xz_end_vec, y_vec, particle intensities), then the 5x5x5
    Gaussian convolve_pencil of the particles in reference tap order; the separable mode
    is compared with it (1e-5 relative);
  * the 64 uneven volumes in one batch plan: every volume's structure;
  * the 1024x1024x256 deconvolution volume: structure, particles and 3 Richardson-Lucy
    iterations in reference order; RL to the 1e-3 stopping rule in both convolution
    modes, whose relative difference is reported;
  * the 256x256x128 volume: the 3x3x3 mean convolve_pencil against the oracle directly.
"""
import json
import os

import numpy as np
import pytest
import xxhash

import libapr_amd as A
import oracle_lib as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX_PATH = os.path.join(HERE, "golden", "large_workloads.json")
FIX = json.load(open(FIX_PATH)) if os.path.exists(FIX_PATH) else {}
CHUNK = 1 << 26


def _apar():
    return A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)


def _digest(a):
    a = np.ascontiguousarray(a)
    return xxhash.xxh3_64_hexdigest(memoryview(a.reshape(-1).view(np.uint8)))


def _check(name, got, rec):
    """got (host array) against a fixture record: digest, else the first differing probe"""
    got = np.asarray(got).reshape(-1)
    assert got.size == rec["size"], f"{name}: size {got.size} != {rec['size']}"
    if _digest(got) == rec["digest"]:
        return
    bad = [(i, got[i].item(), v) for i, v in rec["probes"] if got[i].item() != v]
    raise AssertionError(f"{name}: digest differs; probes differing (index, gpu, oracle): {bad[:5]}")


def _max_rel(got, exp):
    """max |got - exp| / max(|exp|, 1e-3 max|exp|), chunked (float64 per chunk)"""
    scale = 1e-3 * float(np.abs(exp).max()) if exp.size else 1.0
    r = 0.0
    for b in range(0, exp.size, CHUNK):
        e = exp[b:b + CHUNK].astype(np.float64)
        g = got[b:b + CHUNK].astype(np.float64)
        r = max(r, float((np.abs(g - e) / np.maximum(np.abs(e), scale)).max()))
    return r


def _rel_l2(a, b):
    num = den = 0.0
    for s in range(0, b.size, CHUNK):
        x = a[s:s + CHUNK].astype(np.float64)
        y = b[s:s + CHUNK].astype(np.float64)
        num += float(((x - y) ** 2).sum())
        den += float((y ** 2).sum())
    return (num / den) ** 0.5


@pytest.fixture(scope="module")
def lctx():
    c = A.Context(0)
    yield c
    c.close()


def _structure(apr, f):
    assert apr.total == f["total"], (apr.total, f["total"])
    assert (apr.l_min, apr.l_max) == (f["l_min"], f["l_max"])
    lxz, xze, yv = apr.structure()
    assert [int(v) for v in lxz] == f["level_xz_vec"]
    _check("xz_end_vec", xze, f["xz_end_vec"])
    _check("y_vec", yv, f["y_vec"])


@pytest.mark.skipif("large" not in FIX, reason="fixture large_workloads.json has no 'large' entry")
def test_large_volume_get_apr_sample_and_convolution(lctx):
    f = FIX["large"]
    g = FIX["params"]["large"]
    z, x, y = g["shape_zxy"]
    print(f"2048x2048x512: synthesising on the device, expecting P={f['total']}", flush=True)
    img = A.DeviceArray(x * y * z * 2, lctx)
    A.synth_blobs((z, x, y), g["seed"], g["n_blobs"], sigma=g["sigma"], amp=g["amp"], out=img.ptr, ctx=lctx)
    assert _digest(img.to_host(np.uint16)) == f["input"], "device generator differs from the host generator"
    P = f["total"]
    parts = A.DeviceArray(P * 2, lctx)
    conv = A.APRConverter(lctx)
    conv.par = _apar()
    apr, sampled = conv.get_apr_and_sample(img.ptr, parts.ptr, P, dims=(y, x, z))
    img.free()
    print(f"  get_apr P={apr.total} sampled={sampled}", flush=True)
    _structure(apr, f)
    assert sampled
    _check("particles", parts.to_host(np.uint16, P), f["particles"])
    print("  structure and particles bit-exact; 5x5x5 Gaussian convolution", flush=True)
    out = A.DeviceArray(P * 4, lctx)
    st = A.APRStencil.create_gaussian_filter((1.0, 1.0, 1.0), (5, 5, 5), True)
    lctx.set_conv_mode(A.CONV_EXACT)
    A.APRFilter.convolve_pencil(apr, st, parts.ptr, out.ptr, dtype="uint16")
    exact = out.to_host(np.float32, P)
    _check("conv_gauss5", exact, f["conv_gauss5"])
    lctx.set_conv_mode(A.CONV_SEPARABLE)
    try:
        A.APRFilter.convolve_pencil(apr, st, parts.ptr, out.ptr, dtype="uint16")
    finally:
        lctx.set_conv_mode(A.CONV_EXACT)
    r = _max_rel(out.to_host(np.float32, P), exact)
    print(f"  separable vs reference order: max rel {r:.2e}", flush=True)
    assert r < 1e-5
    apr.close()


@pytest.mark.skipif("batch" not in FIX, reason="fixture large_workloads.json has no 'batch' entry")
def test_batch_of_64_uneven_volumes(lctx):
    vols = FIX["batch"]["volumes"]
    g = FIX["params"]["batch"]
    dims = [tuple(v["dims"]) for v in vols]
    batch = A.APRBatch(dims, par=_apar(), ctx=lctx)
    for i, v in enumerate(vols):
        y, x, z = v["dims"]
        A.synth_blobs((z, x, y), v["seed"], v["n_blobs"], sigma=g["sigma"], amp=g["amp"], out=batch.input_address(i),
                      ctx=lctx)
    print(f"batch of {len(vols)} volumes ({batch.voxels / 1e9:.3f} Gvoxels)", flush=True)
    batch.run()
    bad = []
    for i, v in enumerate(vols):
        y, x, z = v["dims"]
        host = np.zeros(y * x * z, np.uint16)
        A._check(A.lib().aprh_memcpy(lctx.h, A.C.c_void_p(host.ctypes.data), batch.input_address(i), host.nbytes))
        assert _digest(host) == v["input"], f"volume {i}: device generator differs from the host generator"
        got = batch.get(i)
        lxz, xze, yv = got.structure()
        ok = (got.total == v["total"] and [int(t) for t in lxz] == v["level_xz_vec"] and _digest(xze) == v["xz_end_vec"]
              and _digest(yv) == v["y_vec"])
        if not ok:
            bad.append((i, v["dims"], got.total, v["total"]))
        got.close()
    batch.close()
    assert not bad, f"volumes differing from the oracle (index, dims, P gpu, P oracle): {bad}"


@pytest.mark.skipif("deconv" not in FIX, reason="fixture large_workloads.json has no 'deconv' entry")
def test_deconvolution_volume_rl(lctx):
    f = FIX["deconv"]
    g = FIX["params"]["deconv"]
    z, x, y = g["shape_zxy"]
    n = x * y * z
    print(f"1024x1024x256 deconvolution volume, expecting P={f['total']}", flush=True)
    gt, img = A.DeviceArray(n * 4, lctx), A.DeviceArray(n * 2, lctx)
    A.synth_deconv((z, x, y), g["seed"], g["n_blobs"], gt.ptr, img.ptr, sigma=g["sigma"], amp=g["amp"],
                   psf_sigma=g["psf_sigma"], psf_taps=g["psf_taps"], ctx=lctx)
    gt.free()
    assert _digest(img.to_host(np.uint16)) == f["input"], "device generator differs from the host generator"
    conv = A.APRConverter(lctx)
    conv.par = _apar()
    apr = conv.get_apr(img.ptr, dims=(y, x, z))
    _structure(apr, f)
    P = apr.total
    parts = A.DeviceArray(P * 2, lctx)
    A.ParticleData().sample_image(apr, img.ptr, out=parts.ptr)
    img.free()
    _check("particles", parts.to_host(np.uint16, P), f["particles"])
    # the PSF exactly as the fixture script built it (its float32 restatement of
    # create_gaussian_filter; the library's expf may differ from it in the last bit)
    psf = O.gaussian_stencil(g["psf_sigma"], (g["psf_taps"],) * 3, True)
    # float particles with fractional values: tree means and reference-order convolution
    pf = (parts.to_host(np.uint16, P).astype(np.float32) * np.float32(1.37)).astype(np.float32)
    d_pf = A.DeviceArray(P * 4, lctx)
    d_pf.from_host(pf)
    tree = A.DeviceArray(max(apr.total_number_tree_particles(), 1) * 4, lctx)
    A.APRTreeNumerics.fill_tree_mean(apr, d_pf.ptr, tree.ptr, dtype="float32")
    _check("tree_mean_f32", tree.to_host(np.float32, apr.total_number_tree_particles()), f["tree_mean_f32"])
    conv_out = A.DeviceArray(P * 4, lctx)
    A.APRFilter.convolve_pencil(apr, psf, d_pf.ptr, conv_out.ptr, normalize=True, dtype="float32")
    _check("conv_psf_f32", conv_out.to_host(np.float32, P), f["conv_psf_f32"])
    for arr in (d_pf, tree, conv_out):
        arr.free()
    u = A.DeviceArray(P * 4, lctx)
    lctx.set_conv_mode(A.CONV_EXACT)
    _, st = A.APRNumerics.richardson_lucy(apr, parts.ptr, u.ptr, psf, g["rl_iters"], normalize=True, dtype="uint16")
    assert st["iterations"] == f["rl"]["iterations"]
    _check("rl_u", u.to_host(np.float32, P), f["rl"]["u"])
    print(f"  structure, particles and {g['rl_iters']} RL iterations bit-exact", flush=True)
    # both convolution modes to the stopping rule
    res = {}
    for mode in (A.CONV_EXACT, A.CONV_SEPARABLE):
        lctx.set_conv_mode(mode)
        _, st = A.APRNumerics.richardson_lucy(apr, parts.ptr, u.ptr, psf, 200, normalize=True, tol=1e-3, dtype="uint16")
        res[mode] = (st["iterations"], u.to_host(np.float32, P))
    lctx.set_conv_mode(A.CONV_EXACT)
    (ie, ue), (is_, us) = res[A.CONV_EXACT], res[A.CONV_SEPARABLE]
    d = _rel_l2(us, ue)
    print(f"  RL to rel update < 1e-3: reference order {ie} iterations, separable {is_}; "
          f"relative L2 difference of the results {d:.2e}", flush=True)
    assert ie == is_ and d < 1e-4
    apr.close()


def test_conversion_volume_mean3_convolution(oracle, lctx):
    """the 256x256x128 volume (500 blobs, seed 0): 3x3x3 mean convolution of its particles,
    bit-equal to the oracle in reference order, 1e-5 relative in separable mode"""
    img = A.synth_blobs((128, 256, 256), 0, 500)
    ref = oracle.get_apr(img, oracle.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10), "oracle", keep=False)
    parts = ref.sample(img)
    st = A.APRStencil.create_mean_filter((3, 3, 3))
    stack = oracle.downsampled_stencils(st, ref.l_max - ref.l_min, False)
    exp = ref.convolve_pencil(stack, parts)
    conv = A.APRConverter(lctx)
    conv.par = _apar()
    apr = conv.get_apr(img)
    assert apr.total == ref.total
    got = A.APRFilter.convolve_pencil(apr, st, A.ParticleData().sample_image(apr, img))
    print(f"256x256x128: P={apr.total}", flush=True)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    lctx.set_conv_mode(A.CONV_SEPARABLE)
    try:
        sep = A.APRFilter.convolve_pencil(apr, st, A.ParticleData().sample_image(apr, img))
    finally:
        lctx.set_conv_mode(A.CONV_EXACT)
    assert _max_rel(sep, exp) < 1e-5
    apr.close()
