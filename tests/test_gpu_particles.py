"""GPU parity of the particle operations against the CPU oracle, through the C-ABI:
tree structure, fill_tree_mean, convolve_pencil / convolve, Richardson-Lucy,
reconstruct_constant, reconstruct_level and reconstruct_smooth.

The GPU sums in the reference's order (tree children: particles then tree nodes,
(z, x, y); convolution taps: (z, x, y)), so every float result is bit-equal to the
oracle, which is itself bit-equal to the reference build (test_oracle_numerics.py).
Larger volumes are checked through size-independent properties: a normalised
stencil maps a constant field to itself, tree means of a constant are constant,
Richardson-Lucy keeps a constant image, and sampling a reconstruction gives the
particles back.
"""
import numpy as np
import pytest

import libapr_amd as A

pytestmark = pytest.mark.gpu

CASES = [((32, 64, 64), 20, 0), ((29, 45, 61), 12, 3), ((16, 16, 130), 8, 4), ((64, 96, 128), 60, 5)]


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _pair(O, ctx, shape, nb, seed):
    """oracle APR and the GPU APR built from the oracle's structure"""
    img = A.synth_blobs(shape, seed, nb)
    o = O.get_apr(img, O.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10), "oracle")
    g = A.APR.from_structure(o.dims, o.xz_end_vec, o.y_vec, ctx)
    print(f"case {shape} seed {seed} P={o.total}", flush=True)
    return img, o, g


def _stencils(O, kind, o):
    if kind == "mean3":
        st = np.full((3, 3, 3), 1.0 / 27, np.float32)
    elif kind == "gauss5":
        st = O.gaussian_stencil((1.0, 1.0, 1.0), (5, 5, 5), True)
    else:
        st = O.gaussian_stencil((1.0, 1.5, 2.0), (3, 5, 7), False)
    return st, O.downsampled_stencils(st, o.l_max - o.l_min, True, "oracle")


@pytest.mark.parametrize("shape,nb,seed", CASES)
def test_tree_and_tree_mean(oracle, ctx, shape, nb, seed):
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    to, tg = o.tree(), g.tree()
    assert tg["total"] == to["total"] and tg["l_max"] == to["l_max"]
    for k in ("level_xz_vec", "xz_end_vec", "y_vec"):
        assert np.array_equal(tg[k], to[k]), k
    parts = o.sample(img)
    assert np.array_equal(_bits(A.APRTreeNumerics.fill_tree_mean(g, parts)), _bits(o.fill_tree_mean(parts)))
    pf = (parts.astype(np.float32) * np.float32(1.37)).astype(np.float32)
    assert np.array_equal(_bits(A.APRTreeNumerics.fill_tree_mean(g, pf)), _bits(o.fill_tree_mean(pf)))


@pytest.mark.parametrize("shape,nb,seed", CASES)
@pytest.mark.parametrize("kind", ["mean3", "gauss5", "aniso"])
@pytest.mark.parametrize("reflect", [True, False])
def test_convolve_pencil(oracle, ctx, shape, nb, seed, kind, reflect):
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    parts = o.sample(img)
    st, stack = _stencils(oracle, kind, o)
    pf = parts.astype(np.float32)
    exp = o.convolve_pencil(stack, pf, reflect)
    got = A.APRFilter.convolve_pencil(g, stack, pf, reflect_boundary=reflect)
    nbad = int((_bits(got) != _bits(exp)).sum())
    assert nbad == 0, f"{nbad} of {got.size} differ; max abs {np.abs(got - exp).max()}"
    # u16 input, and the single-stencil form that prepares the stack itself
    assert np.array_equal(_bits(A.APRFilter.convolve(g, stack, parts, reflect_boundary=reflect)),
                          _bits(o.convolve_pencil(stack, parts, reflect)))
    assert np.array_equal(_bits(A.APRFilter.convolve_pencil(g, st, pf, reflect_boundary=reflect, normalize=True)), _bits(exp))


def test_convolve_unit_stencil_is_identity(oracle, ctx):
    img, o, g = _pair(oracle, ctx, *CASES[0])
    parts = o.sample(img).astype(np.float32)
    one = np.zeros((3, 3, 3), np.float32)
    one[1, 1, 1] = 1
    out = A.APRFilter.convolve_pencil(g, [one], parts)
    assert np.array_equal(out, parts)


@pytest.mark.parametrize("shape,nb,seed", CASES[:2])
def test_richardson_lucy(oracle, ctx, shape, nb, seed):
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    parts = o.sample(img)
    psf = oracle.gaussian_stencil((1.0, 1.0, 2.0), (9, 9, 9), True)
    uo, it, rel_o = o.richardson_lucy(parts, psf, 3)
    ug, st = A.APRNumerics.richardson_lucy(g, parts, None, psf, 3, normalize=True)
    assert st["iterations"] == it == 3
    assert np.array_equal(_bits(ug), _bits(uo))
    assert abs(st["last_rel_update"] - rel_o) <= 1e-9 * max(rel_o, 1e-30)
    # resume continues from the estimate (APRNumerics.hpp:559-562)
    uo2, _, _ = o.richardson_lucy(parts, psf, 2, u=uo)
    ug2, _ = A.APRNumerics.richardson_lucy(g, parts, ug.copy(), psf, 2, normalize=True, resume=True)
    assert np.array_equal(_bits(ug2), _bits(uo2))
    # float input
    pf = parts.astype(np.float32)
    uo3, _, _ = o.richardson_lucy(pf, psf, 2)
    ug3, _ = A.APRNumerics.richardson_lucy(g, pf, None, psf, 2, normalize=True)
    assert np.array_equal(_bits(ug3), _bits(uo3))


@pytest.mark.parametrize("shape,nb,seed", CASES)
def test_richardson_lucy_tv(oracle, ctx, shape, nb, seed):
    """APRNumerics::richardson_lucy_tv: bit-equal to the oracle (whose restatement is bit-equal
    to the reference build), u16 and f32 input, fresh and resumed; reg = 0 is plain RL"""
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    parts = o.sample(img)
    psf = oracle.gaussian_stencil((1.0, 1.0, 1.5), (5, 5, 5), True)
    print(f"RL-TV {shape}: P={o.total}", flush=True)
    uo = o.richardson_lucy_tv(parts, psf, 3, 0.01)
    ug = A.APRNumerics.richardson_lucy_tv(g, parts, None, psf, 3, 0.01, normalize=True)
    assert np.array_equal(_bits(ug), _bits(uo))
    pf = parts.astype(np.float32) * np.float32(0.7)
    uo2 = o.richardson_lucy_tv(pf, psf, 2, 0.02, u=uo)
    ug2 = A.APRNumerics.richardson_lucy_tv(g, pf, ug.copy(), psf, 2, 0.02, normalize=True, resume=True)
    assert np.array_equal(_bits(ug2), _bits(uo2))
    plain, _ = A.APRNumerics.richardson_lucy(g, parts, None, psf, 2, normalize=True)
    assert np.array_equal(_bits(A.APRNumerics.richardson_lucy_tv(g, parts, None, psf, 2, 0.0, normalize=True)), _bits(plain))


def test_richardson_lucy_stopping_rule(oracle, ctx):
    img, o, g = _pair(oracle, ctx, *CASES[0])
    parts = o.sample(img)
    psf = oracle.gaussian_stencil((1.0, 1.0, 2.0), (9, 9, 9), True)
    uo, it, rel = o.richardson_lucy(parts, psf, 200, tol=1e-3)
    ug, st = A.APRNumerics.richardson_lucy(g, parts, None, psf, 200, normalize=True, tol=1e-3)
    assert st["iterations"] == it < 200 and st["last_rel_update"] < 1e-3
    assert np.array_equal(_bits(ug), _bits(uo))


@pytest.mark.parametrize("shape,nb,seed", CASES[:3])
@pytest.mark.parametrize("delta", [0, -1, -2])
def test_reconstruct_constant(oracle, ctx, shape, nb, seed, delta):
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    parts = o.sample(img)
    for out_dtype in (np.uint16, np.float32):
        assert np.array_equal(A.APRReconstruction.reconstruct_constant(g, parts, delta, out_dtype),
                              o.reconstruct(parts, delta, out_dtype))
    pf = (parts * np.float32(0.73)).astype(np.float32)
    for out_dtype in (np.uint16, np.float32):
        got = A.APRReconstruction.reconstruct_constant(g, pf, delta, out_dtype)
        exp = o.reconstruct(pf, delta, out_dtype)
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


def test_gpu_apr_particle_ops(oracle, ctx):
    """the same operations on an APR made by the GPU converter"""
    img = A.synth_blobs((61, 63, 65), 1, 30)
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)
    g = conv.get_apr(img)
    o = oracle.get_apr(img, oracle.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10), "oracle")
    parts = A.ParticleData().sample_image(g, img)
    _, stack = _stencils(oracle, "gauss5", o)
    assert np.array_equal(_bits(A.APRFilter.convolve_pencil(g, stack, parts)), _bits(o.convolve_pencil(stack, parts, True)))
    assert np.array_equal(A.APRReconstruction.reconstruct_constant(g, parts), o.reconstruct(parts))


def test_malformed_structure_rejected(oracle, ctx):
    img = A.synth_blobs((32, 64, 64), 0, 20)
    o = oracle.get_apr(img, oracle.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10), "oracle")
    y = o.y_vec.copy()
    y[len(y) // 2] = 65000            # beyond y_num of any level
    with pytest.raises(A.AprhError):
        A.APR.from_structure(o.dims, o.xz_end_vec, y, ctx)
    x = o.xz_end_vec.copy()
    x[-1] += 1                        # total mismatch
    with pytest.raises(A.AprhError):
        A.APR.from_structure(o.dims, x, o.y_vec, ctx)
    with pytest.raises(A.AprhError):
        A.APR.from_structure(o.dims, o.xz_end_vec[:-1], o.y_vec, ctx)


def test_argument_errors(oracle, ctx):
    img, o, g = _pair(oracle, ctx, *CASES[0])
    parts = o.sample(img)
    with pytest.raises(A.AprhError):
        A.APRFilter.convolve_pencil(g, [np.ones((2, 3, 3), np.float32)], parts)          # even stencil
    with pytest.raises(A.AprhError):
        A.APRReconstruction.reconstruct_constant(g, parts, -(o.l_max))                # below l_min
    with pytest.raises(A.AprhError):
        A.APRNumerics.richardson_lucy(g, parts, None, np.ones((3, 3, 3), np.float32), -1)


def test_properties_at_size(ctx):
    """size-independent checks on a 512x512x256 volume (GPU only)"""
    shape = (256, 512, 512)
    img = A.synth_blobs(shape, 7, 3000, sigma=(1.5, 6.0), amp=(300.0, 4000.0))
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)
    g = conv.get_apr(img)
    parts = A.ParticleData().sample_image(g, img)
    print(f"P={g.total}", flush=True)
    # reconstruction then sampling returns the particles
    rec = A.APRReconstruction.reconstruct_constant(g, parts)
    assert np.array_equal(A.ParticleData().sample_image(g, rec), parts)
    # constant field: tree means, normalised convolution and RL keep it
    c = np.full(g.total, 7.0, np.float32)
    tm = A.APRTreeNumerics.fill_tree_mean(g, c)
    np.testing.assert_allclose(tm, 7.0, rtol=1e-6)
    st = A.APRStencil.create_gaussian_filter((1.0, 1.0, 1.0), (5, 5, 5), True)
    out = A.APRFilter.convolve_pencil(g, st, c, normalize=True)
    np.testing.assert_allclose(out, 7.0, rtol=1e-5)
    # u = 1 -> blur 1 -> ratio 7 -> u = 7, then ratio 1 keeps it
    u, s = A.APRNumerics.richardson_lucy(g, c, None, st, 3, normalize=True)
    np.testing.assert_allclose(u, 7.0, rtol=1e-4)
    assert s["iterations"] == 3


@pytest.fixture(scope="module")
def sep_ctx():
    c = A.Context(0)
    c.set_conv_mode(A.CONV_SEPARABLE)
    yield c
    c.close()


def _rel(got, exp):
    exp = exp.astype(np.float64)
    return np.abs(got.astype(np.float64) - exp) / np.maximum(np.abs(exp), 1e-3 * np.abs(exp).max())


@pytest.mark.parametrize("shape,nb,seed", CASES)
@pytest.mark.parametrize("kind", ["mean3", "gauss5", "aniso"])
@pytest.mark.parametrize("reflect", [True, False])
def test_convolve_separable_mode(oracle, sep_ctx, shape, nb, seed, kind, reflect):
    """CONV_SEPARABLE: rank-1 stencils as three 1-D passes; within 1e-5 relative of the
    reference-order result, and as close as it to the float64 ground truth"""
    img, o, g = _pair(oracle, sep_ctx, shape, nb, seed)
    parts = o.sample(img).astype(np.float32)
    st, stack = _stencils(oracle, kind, o)
    exp = o.convolve_pencil(stack, parts, reflect)
    got = A.APRFilter.convolve_pencil(g, stack, parts, reflect_boundary=reflect)
    r = _rel(got, exp).max()
    print(f"separable vs reference order: max rel {r:.2e}")
    assert r < 1e-5
    if shape == CASES[1][0]:
        from test_oracle_numerics import conv_gt
        t = o.tree()
        gt = conv_gt(o, stack, parts.astype(np.float64), o.fill_tree_mean(parts).astype(np.float64), t, reflect)
        e_exact, e_sep = _rel(exp, gt).max(), _rel(got, gt).max()
        print(f"vs float64 ground truth: reference order {e_exact:.2e}, separable {e_sep:.2e}")
        assert e_sep < 1e-5 and e_exact < 1e-5


def test_richardson_lucy_tv_separable_mode(oracle, sep_ctx):
    img, o, g = _pair(oracle, sep_ctx, *CASES[0])
    parts = o.sample(img)
    psf = oracle.gaussian_stencil((1.0, 1.0, 2.0), (9, 9, 9), True)
    uo = o.richardson_lucy_tv(parts, psf, 3, 0.01)
    ug = A.APRNumerics.richardson_lucy_tv(g, parts, None, psf, 3, 0.01, normalize=True)
    r = _rel(ug, uo).max()
    l2 = float(np.linalg.norm(ug.astype(np.float64) - uo) / np.linalg.norm(uo.astype(np.float64)))
    print(f"RL-TV 3 iterations, separable vs reference order: max rel {r:.2e}, rel L2 {l2:.2e}")
    # the normalised gradient turns where |grad u| is near zero, so single particles may differ
    # more than the convolutions do; the whole estimate stays within float rounding
    assert l2 < 1e-5


def test_richardson_lucy_separable_mode(oracle, sep_ctx):
    img, o, g = _pair(oracle, sep_ctx, *CASES[0])
    parts = o.sample(img)
    psf = oracle.gaussian_stencil((1.0, 1.0, 2.0), (9, 9, 9), True)
    uo, _, _ = o.richardson_lucy(parts, psf, 3)
    ug, st = A.APRNumerics.richardson_lucy(g, parts, None, psf, 3, normalize=True)
    r = _rel(ug, uo).max()
    print(f"RL 3 iterations, separable vs reference order: max rel {r:.2e}")
    assert st["iterations"] == 3 and r < 1e-5


def _level_image(o):
    """reconstruct_level restated from the structure (APRReconstruction.hpp:131-171): the
    particle of level l at row (z, x), position y covers the block [z, x, y] * 2^(l_max - l)"""
    y, x, z = o.dims
    img = np.zeros((z, x, y), np.uint16)
    lxz, xze, yv = o.level_xz_vec, o.xz_end_vec, o.y_vec
    for lev in range(o.l_min, o.l_max + 1):
        s = 2 ** (o.l_max - lev)
        xl, zl = -(-x // s), -(-z // s)
        for zi in range(zl):
            for xi in range(xl):
                r = int(lxz[lev]) + zi * xl + xi
                b = int(xze[r - 1]) if r > 0 else 0
                for yy in yv[b:int(xze[r])]:
                    img[zi * s:(zi + 1) * s, xi * s:(xi + 1) * s, int(yy) * s:(int(yy) + 1) * s] = lev
    return img


@pytest.mark.parametrize("shape,nb,seed", CASES[:3])
def test_reconstruct_level(oracle, ctx, shape, nb, seed):
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    expect = _level_image(o)
    # the same image as the constant reconstruction of particles valued by their levels
    lev = np.zeros(o.total, np.uint16)
    for l in range(o.l_min, o.l_max + 1):
        r0 = int(o.level_xz_vec[l])
        b = int(o.xz_end_vec[r0 - 1]) if r0 > 0 else 0
        e = int(o.xz_end_vec[int(o.level_xz_vec[l + 1]) - 1])
        lev[b:e] = l
    assert np.array_equal(o.reconstruct(lev, 0, np.uint16), expect)
    assert int(expect.min()) >= o.l_min and int(expect.max()) == o.l_max
    for out_dtype in (np.uint16, np.float32):
        got = A.APRReconstruction.reconstruct_level(g, out_dtype)
        assert got.dtype == out_dtype and np.array_equal(got, expect.astype(out_dtype))


@pytest.mark.parametrize("shape,nb,seed", CASES[:3] + [((3, 9, 40), 2, 8)])
def test_reconstruct_smooth(oracle, ctx, shape, nb, seed):
    """reconstruct_smooth bit-equal to the oracle: u16 and f32 particles into u16 and f32 images,
    the default scale and an anisotropic one, and a scale so large that no pixel is smoothed"""
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    pu = o.sample(img)
    pf = (pu.astype(np.float32) * np.float32(0.37)) + np.float32(1.5)
    for parts in (pu, pf):
        for out_dtype in (np.uint16, np.float32):
            for sc in ((2.0, 2.0, 2.0), (1.0, 0.7, 3.0), (1e6, 1e6, 1e6)):
                exp = o.reconstruct_smooth(parts, scale=sc, out_dtype=out_dtype)
                got = A.APRReconstruction.reconstruct_smooth(g, parts, scale_d=sc, out_dtype=out_dtype)
                assert got.dtype == exp.dtype
                if out_dtype == np.float32:
                    ok = np.array_equal(_bits(got), _bits(exp))
                else:
                    ok = np.array_equal(got, exp)
                assert ok, (parts.dtype, out_dtype, sc, int((got != exp).sum()))


def _patch_restated(o, values, tree_values, p, out_dtype):
    """the patch loop of APRReconstruction.hpp:296-394 (reconstruct_level :397-490 is the same
    loop with level-valued particles and the patch level for tree nodes), row by row"""
    y, x, z = o.dims
    max_level = o.l_max + p.level_delta
    img = np.zeros((p.z_end - p.z_begin, p.x_end - p.x_begin, p.y_end - p.y_begin), out_dtype)
    lxz, xze, yv = o.level_xz_vec, o.xz_end_vec, o.y_vec

    def row(lxz_, xze_, lev, zz, xx):
        xl = -(-x // 2 ** (o.l_max - lev))
        r = int(lxz_[lev]) + zz * xl + xx
        return (int(xze_[r - 1]) if r > 0 else 0), int(xze_[r])

    for xx in range(p.x_begin, p.x_end):
        for zz in range(p.z_begin, p.z_end):
            for lev in range(min(max_level, o.l_max), o.l_min - 1, -1):
                ylev = -(-y // 2 ** (o.l_max - lev))
                if lev == max_level:
                    b, e = row(lxz, xze, lev, zz, xx)
                    for i in range(b, e):
                        if p.y_begin <= yv[i] < p.y_end:
                            img[zz - p.z_begin, xx - p.x_begin, int(yv[i]) - p.y_begin] = values[i]
                else:
                    step = 2 ** (max_level - lev)
                    b, e = row(lxz, xze, lev, zz // step, xx // step)
                    ybl, yel = p.y_begin // step, min(-(-p.y_end // step), ylev)
                    for i in range(b, e):
                        if ybl <= yv[i] < yel:
                            lo = max(int(yv[i]) * step, p.y_begin) - p.y_begin
                            hi = min((int(yv[i]) + 1) * step, p.y_end) - p.y_begin
                            img[zz - p.z_begin, xx - p.x_begin, lo:hi] = values[i]
    if max_level < o.l_max:
        t = o.tree()
        for xx in range(p.x_begin, p.x_end):
            for zz in range(p.z_begin, p.z_end):
                b, e = row(t["level_xz_vec"], t["xz_end_vec"], max_level, zz, xx)
                for i in range(b, e):
                    yy = int(t["y_vec"][i])
                    if p.y_begin <= yy < p.y_end:
                        img[zz - p.z_begin, xx - p.x_begin, yy - p.y_begin] = tree_values[i]
    return img


def _levels_of(o):
    lev = np.zeros(o.total, np.uint16)
    for l in range(o.l_min, o.l_max + 1):
        r0 = int(o.level_xz_vec[l])
        b = int(o.xz_end_vec[r0 - 1]) if r0 > 0 else 0
        lev[b:int(o.xz_end_vec[int(o.level_xz_vec[l + 1]) - 1])] = l
    return lev


PATCHES = [dict(x_begin=3, x_end=40, y_begin=5, y_end=50, z_begin=2, z_end=20),
           dict(x_begin=-4, x_end=1000, y_begin=17, y_end=-1, z_begin=9, z_end=-1),
           dict(x_begin=1, x_end=12, y_begin=3, y_end=21, z_begin=0, z_end=7, level_delta=-1),
           dict(x_begin=2, x_end=9, y_begin=0, y_end=-1, z_begin=1, z_end=5, level_delta=-2),
           dict(x_begin=7, x_end=50, y_begin=11, y_end=90, z_begin=3, z_end=19, level_delta=1)]


@pytest.mark.parametrize("shape,nb,seed", CASES[:2])
@pytest.mark.parametrize("pi", range(len(PATCHES)))
def test_reconstruct_patch(oracle, ctx, shape, nb, seed, pi):
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    parts = o.sample(img)
    p = A.ReconPatch(**PATCHES[pi])
    assert p.check_limits(g)
    d = p.level_delta
    # the restated loop against a cut of the oracle's whole-image reconstruction
    full = o.reconstruct(parts, min(d, 0), np.uint16)
    if d > 0:
        full = full.repeat(2 ** d, 0).repeat(2 ** d, 1).repeat(2 ** d, 2)
    cut = full[p.z_begin:p.z_end, p.x_begin:p.x_end, p.y_begin:p.y_end]
    tree = o.fill_tree_mean(parts) if d < 0 else None
    exp = _patch_restated(o, parts, tree, p, np.uint16)
    assert np.array_equal(exp, cut)
    got = A.APRReconstruction.reconstruct_constant(g, parts, patch=A.ReconPatch(**PATCHES[pi]))
    assert got.dtype == np.uint16 and np.array_equal(got, exp)
    gotf = A.APRReconstruction.reconstruct_constant(g, parts, out_dtype=np.float32, patch=A.ReconPatch(**PATCHES[pi]))
    expf = _patch_restated(o, parts.astype(np.float32), tree, p, np.float32)
    assert np.array_equal(_bits(gotf), _bits(expf))
    # levels: the particles' levels, capped at the patch level (tree nodes take the patch level)
    lev = _levels_of(o)
    maxl = o.l_max + d
    tl = np.full(o.tree()["total"], maxl, np.uint16) if d < 0 else None
    expl = _patch_restated(o, np.minimum(lev, maxl), tl, p, np.uint16)
    gotl = A.APRReconstruction.reconstruct_level(g, patch=A.ReconPatch(**PATCHES[pi]))
    assert np.array_equal(gotl, expl)
    if d == 0:
        assert np.array_equal(expl, _level_image(o)[p.z_begin:p.z_end, p.x_begin:p.x_end, p.y_begin:p.y_end])


def test_reconstruct_patch_limits(oracle, ctx):
    img, o, g = _pair(oracle, ctx, *CASES[1])
    y, x, z = o.dims
    p = A.ReconPatch(x_begin=-3, y_end=10 ** 6, z_begin=2, level_delta=-1)
    assert p.check_limits(g)
    assert (p.x_begin, p.x_end, p.y_begin, p.y_end, p.z_begin, p.z_end) == (0, -(-x // 2), 0, -(-y // 2), 2, -(-z // 2))
    for bad in (dict(x_begin=5, x_end=5), dict(y_begin=10 ** 6), dict(z_begin=3, z_end=2), dict(level_delta=40)):
        assert not A.ReconPatch(**bad).check_limits(g)
        with pytest.raises(ValueError):
            A.APRReconstruction.reconstruct_level(g, patch=A.ReconPatch(**bad))
    with pytest.raises(A.AprhError):
        A.APRReconstruction.reconstruct_level(g, patch=A.ReconPatch(level_delta=-(o.l_max - o.l_min) - 1))


def _gradient_stack(o, dim, delta, sobel):
    """APRNumerics::gradient_cfd / gradient_sobel's stencils (APRNumerics.hpp:165-213,
    APRStencil.hpp:135-152, :254-297) restated in float32, [z, x, y]"""
    y, x, z = o.dims
    nd = int(y > 1) + int(x > 1) + int(z > 1)
    h = np.float32(1.0) / (np.float32(2.0) * np.float32(delta))
    diff = np.array([-h, 0.0, h], np.float32)
    sm = np.array([0.25, 0.5, 0.25], np.float32)
    fy, fx, fz = (diff if dim == 0 else sm), (diff if dim == 1 else sm), (diff if dim == 2 else sm)
    if sobel and nd == 3:
        st = (fy[None, None, :] * fx[None, :, None]) * fz[:, None, None]
    elif sobel and nd == 2 and dim != 2:
        st = (fy[None, None, :] * fx[None, :, None]).astype(np.float32)
    else:
        st = diff.reshape([(3, 1, 1), (1, 3, 1), (1, 1, 3)][2 - dim])
    st = st.astype(np.float32)
    return [st / np.float32(2.0 ** d) for d in range(max(o.l_max - o.l_min, 1))]


@pytest.mark.parametrize("shape,nb,seed", CASES[:3])
@pytest.mark.parametrize("sobel", [False, True])
def test_gradients(oracle, ctx, shape, nb, seed, sobel):
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    parts = o.sample(img)
    pf = (parts.astype(np.float32) * np.float32(0.731)).astype(np.float32)
    fn = A.APRNumerics.gradient_sobel if sobel else A.APRNumerics.gradient_cfd
    grads = []
    for dim, delta in ((0, 1.0), (1, 0.7), (2, 2.5)):
        stack = _gradient_stack(o, dim, delta, sobel)
        exp = o.convolve_pencil(stack, parts, True)
        got = fn(g, parts, dimension=dim, delta=delta)
        assert np.array_equal(_bits(got), _bits(exp)), (dim, int((_bits(got) != _bits(exp)).sum()))
        assert np.array_equal(_bits(fn(g, pf, dimension=dim, delta=delta)), _bits(o.convolve_pencil(stack, pf, True)))
        grads.append(exp)
    # the magnitude in the reference's order: g_y^2, + g_x^2, + g_z^2, sqrtf
    m = grads[0] * grads[0]
    m = (m + grads[1] * grads[1]).astype(np.float32)
    m = (m + grads[2] * grads[2]).astype(np.float32)
    m = np.sqrt(m).astype(np.float32)
    mfn = A.APRNumerics.gradient_magnitude_sobel if sobel else A.APRNumerics.gradient_magnitude_cfd
    assert np.array_equal(_bits(mfn(g, parts, deltas=(1.0, 0.7, 2.5))), _bits(m))
    # and against the reference's own APRNumerics, where its build is present
    if oracle.load("ref") is not None and hasattr(oracle.load("ref"), "orc_gradient_u16"):
        r = oracle.apr_from_structure(o.dims, o.xz_end_vec, o.y_vec, "ref")
        for dim, delta in ((0, 1.0), (1, 0.7), (2, 2.5)):
            for p in (parts, pf):
                assert np.array_equal(_bits(fn(g, p, dimension=dim, delta=delta)), _bits(r.gradient(p, sobel, dim, delta)))
        assert np.array_equal(_bits(mfn(g, pf, deltas=(1.0, 0.7, 2.5))), _bits(r.gradient_magnitude(pf, sobel, (1.0, 0.7, 2.5))))
    with pytest.raises(A.AprhError):
        fn(g, parts, dimension=3)


def test_gradient_of_a_constant_is_zero(oracle, ctx):
    img, o, g = _pair(oracle, ctx, *CASES[3])
    c = np.full(o.total, 1234, np.uint16)
    for fn in (A.APRNumerics.gradient_cfd, A.APRNumerics.gradient_sobel):
        for dim in range(3):
            assert np.all(fn(g, c, dimension=dim) == 0.0)
    assert np.all(A.APRNumerics.gradient_magnitude_sobel(g, c) == 0.0)


@pytest.mark.parametrize("shape,nb,seed", CASES[:3])
@pytest.mark.parametrize("size", [(3, 3, 3), (5, 3, 1)])
def test_local_std(oracle, ctx, shape, nb, seed, size):
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    parts = o.sample(img)
    assert np.array_equal(_bits(A.APRNumerics.local_std(g, parts, size=size)), _bits(o.local_std(parts, size)))
    pf = (parts.astype(np.float32) * np.float32(0.37)).astype(np.float32)
    assert np.array_equal(_bits(A.APRNumerics.local_std(g, pf, size=size)), _bits(o.local_std(pf, size)))


@pytest.mark.parametrize("shape,nb,seed", CASES + [((1, 70, 90), 6, 5), ((33, 17, 9), 5, 6)])
def test_face_neighbour_filters(oracle, ctx, shape, nb, seed):
    """APRNumerics::face_neighbour_filter / seperable_face_neighbour_filter, bit-equal to the
    oracle (itself bit-equal to the reference build), u16 and f32 input"""
    img, o, g = _pair(oracle, ctx, shape, nb, seed)
    parts = o.sample(img)
    pf = (parts.astype(np.float32) * np.float32(0.37)).astype(np.float32)
    f = np.array([0.25, 0.5, 0.3], np.float32)
    for p in (parts, pf):
        for dim in range(3):
            got = A.APRNumerics.face_neighbour_filter(g, p, f, dim)
            exp = o.face_neighbour_filter(p.astype(np.float32), f, dim)
            assert np.array_equal(_bits(got), _bits(exp)), (dim, int((_bits(got) != _bits(exp)).sum()))
    for rep in (0, 1, 2):
        got = A.APRNumerics.seperable_face_neighbour_filter(g, pf, f, rep)
        assert np.array_equal(_bits(got), _bits(o.face_neighbour_filter(pf, f, 0, rep))), rep

