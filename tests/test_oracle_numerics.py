"""Pins the oracle's particle operations (tree, fill_tree_mean, downsampled
stencils, convolve_pencil / convolve, richardson_lucy) -- CPU only.

* bit equality with the reference build (oracle/_ref) on synthetic APRs;
* the reference test's ground truth for convolution (test/FilterTestHelpers.hpp:
  98-192): reconstruct each level piecewise-constant (coarser particles fill their
  blocks, finer regions take the tree means) and apply the stencil densely with
  reflective or zero boundaries, in float64 -- tolerance 1e-5 relative (the
  reference's compareParticles allows 1e-4, TestTools.hpp:93-111);
* closed forms: a constant input gives constant tree means; normalised stencils
  sum to 1; Richardson-Lucy keeps a constant image constant.
"""
import numpy as np
import pytest

import libapr_amd as A
import oracle_lib as O

CASES = [((32, 64, 64), 20, 0), ((29, 45, 61), 12, 3), ((16, 16, 130), 8, 4)]


def make_apr(shape, nb, seed, which="oracle"):
    img = A.synth_blobs(shape, seed, nb)
    par = O.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)
    return img, O.get_apr(img, par, which)


def need_ref():
    if O.load("ref") is None:
        pytest.skip("reference build (oracle/_ref) not available")


def particle_coords(apr):
    """(level, z, x, y) of every particle, in particle order"""
    lv, zz, xx = np.zeros(apr.total, np.int64), np.zeros(apr.total, np.int64), np.zeros(apr.total, np.int64)
    y, x, z = apr.dims
    lxz = apr.level_xz_vec.astype(np.int64)
    xze = apr.xz_end_vec.astype(np.int64)
    for l in range(apr.l_min, apr.l_max + 1):
        s = 2 ** (apr.l_max - l)
        xn, zn = -(-x // s), -(-z // s)
        rows = np.arange(xn * zn)
        b = xze[lxz[l] + rows - 1]
        e = xze[lxz[l] + rows]
        for r in np.nonzero(e > b)[0]:
            lv[b[r]:e[r]] = l
            zz[b[r]:e[r]] = r // xn
            xx[b[r]:e[r]] = r % xn
    return lv, zz, xx, apr.y_vec.astype(np.int64)


def tree_coords(apr, t):
    y, x, z = apr.dims
    n = t["total"]
    lv, zz, xx = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n, np.int64)
    lxz, xze = t["level_xz_vec"].astype(np.int64), t["xz_end_vec"].astype(np.int64)
    for l in range(t["l_min"], t["l_max"] + 1):
        s = 2 ** (apr.l_max - l)
        xn, zn = -(-x // s), -(-z // s)
        rows = np.arange(xn * zn)
        b, e = xze[lxz[l] + rows - 1], xze[lxz[l] + rows]
        for r in np.nonzero(e > b)[0]:
            lv[b[r]:e[r]] = l
            zz[b[r]:e[r]] = r // xn
            xx[b[r]:e[r]] = r % xn
    return lv, zz, xx, t["y_vec"].astype(np.int64)


def recon_level(apr, parts, tree, t, level):
    """piecewise-constant reconstruction at `level` (APRReconstruction.hpp:297-394)"""
    y, x, z = apr.dims
    s = 2 ** (apr.l_max - level)
    dims = (-(-z // s), -(-x // s), -(-y // s))
    img = np.full(dims, np.nan)
    lv, pz, px, py = particle_coords(apr)
    for l in range(apr.l_min, level + 1):
        m = lv == l
        st = 2 ** (level - l)
        for (a, b, c, v) in zip(pz[m], px[m], py[m], parts[m]):
            img[a * st:(a + 1) * st, b * st:(b + 1) * st, c * st:(c + 1) * st] = v
    if level < apr.l_max:
        tl, tz, tx, ty = tree_coords(apr, t)
        m = tl == level
        img[tz[m], tx[m], ty[m]] = tree[m]
    return img


def conv_gt(apr, stencils, parts, tree, t, reflect):
    """FilterTestHelpers::compute_convolution_gt in float64"""
    out = np.zeros(apr.total)
    lv, pz, px, py = particle_coords(apr)
    for li, level in enumerate(range(apr.l_max, apr.l_min - 1, -1)):
        st = stencils[min(li, len(stencils) - 1)].astype(np.float64)
        hz, hx, hy = [(n - 1) // 2 for n in st.shape]
        img = recon_level(apr, parts, tree, t, level)
        Z, X, Y = img.shape
        if reflect:
            pad = np.pad(img, ((hz, hz), (hx, hx), (hy, hy)), mode="reflect")
        else:
            pad = np.pad(img, ((hz, hz), (hx, hx), (hy, hy)), mode="constant")
        m = np.nonzero(lv == level)[0]
        for i in m:
            blk = pad[pz[i]:pz[i] + 2 * hz + 1, px[i]:px[i] + 2 * hx + 1, py[i]:py[i] + 2 * hy + 1]
            out[i] = float((blk * st).sum())
    return out


def stencil_stack(kind, apr, normalize):
    if kind == "mean3":
        st = np.full((3, 3, 3), 1.0 / 27, np.float32)
    elif kind == "gauss5":
        st = O.gaussian_stencil((1.0, 1.0, 1.0), (5, 5, 5), True)
    else:
        st = O.gaussian_stencil((1.0, 1.5, 2.0), (3, 5, 7), False)
    return O.downsampled_stencils(st, apr.l_max - apr.l_min, normalize, "oracle")


@pytest.mark.parametrize("shape,nb,seed", CASES)
def test_tree_and_tree_mean_equal_reference(shape, nb, seed):
    need_ref()
    img, o = make_apr(shape, nb, seed, "oracle")
    _, r = make_apr(shape, nb, seed, "ref")
    to, tr = o.tree(), r.tree()
    assert to["total"] == tr["total"]
    for k in ("level_xz_vec", "xz_end_vec", "y_vec"):
        assert np.array_equal(to[k], tr[k])
    parts = o.sample(img)
    assert np.array_equal(o.fill_tree_mean(parts).view(np.uint32), r.fill_tree_mean(parts).view(np.uint32))
    pf = (parts.astype(np.float32) * np.float32(1.37)).astype(np.float32)
    assert np.array_equal(o.fill_tree_mean(pf).view(np.uint32), r.fill_tree_mean(pf).view(np.uint32))


@pytest.mark.parametrize("shape,nb,seed", CASES[:2])
def test_tree_mean_constant(shape, nb, seed):
    """boundary scale factors make a constant input give a constant tree"""
    img, o = make_apr(shape, nb, seed)
    tm = o.fill_tree_mean(np.full(o.total, 7.0, np.float32))
    np.testing.assert_allclose(tm, 7.0, rtol=1e-6)


@pytest.mark.parametrize("kind", ["mean3", "gauss5", "aniso"])
@pytest.mark.parametrize("normalize", [False, True])
def test_stencils_equal_reference(kind, normalize):
    need_ref()
    img, o = make_apr((32, 64, 64), 20, 0)
    if kind == "mean3":
        st = np.full((3, 3, 3), 1.0 / 27, np.float32)
    elif kind == "gauss5":
        st = O.gaussian_stencil((1.0, 1.0, 1.0), (5, 5, 5), True)
    else:
        st = O.gaussian_stencil((1.0, 1.5, 2.0), (3, 5, 7), False)
    so = O.downsampled_stencils(st, 6, normalize, "oracle")
    sr = O.downsampled_stencils(st, 6, normalize, "ref")
    for a, b in zip(so, sr):
        assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))
        if normalize:
            assert abs(float(a.astype(np.float64).sum()) - 1.0) < 1e-5


@pytest.mark.parametrize("shape,nb,seed", CASES)
@pytest.mark.parametrize("kind", ["mean3", "gauss5", "aniso"])
@pytest.mark.parametrize("reflect", [True, False])
def test_convolve_equal_reference(shape, nb, seed, kind, reflect):
    need_ref()
    img, o = make_apr(shape, nb, seed, "oracle")
    _, r = make_apr(shape, nb, seed, "ref")
    parts = o.sample(img)
    st = stencil_stack(kind, o, True)
    pf = parts.astype(np.float32)
    co = o.convolve_pencil(st, pf, reflect)
    assert np.array_equal(co.view(np.uint32), r.convolve_pencil(st, pf, reflect).view(np.uint32))
    assert np.array_equal(co.view(np.uint32), r.convolve_pencil(st, pf, reflect, which="full").view(np.uint32))
    assert np.array_equal(o.convolve_pencil(st, parts, reflect).view(np.uint32),
                          r.convolve_pencil(st, parts, reflect).view(np.uint32))


@pytest.mark.parametrize("kind", ["mean3", "gauss5", "aniso"])
@pytest.mark.parametrize("reflect", [True, False])
def test_convolve_ground_truth(kind, reflect):
    img, o = make_apr((29, 45, 61), 12, 3)
    parts = o.sample(img).astype(np.float32)
    st = stencil_stack(kind, o, True)
    out = o.convolve_pencil(st, parts, reflect)
    t = o.tree()
    gt = conv_gt(o, st, parts.astype(np.float64), o.fill_tree_mean(parts).astype(np.float64), t, reflect)
    rel = np.abs(out - gt) / np.maximum(np.abs(gt), 1e-3)
    assert rel.max() < 1e-5, rel.max()


@pytest.mark.parametrize("shape,nb,seed", CASES[:2])
def test_richardson_lucy_equal_reference(shape, nb, seed):
    need_ref()
    img, o = make_apr(shape, nb, seed, "oracle")
    _, r = make_apr(shape, nb, seed, "ref")
    parts = o.sample(img)
    psf = O.gaussian_stencil((1.0, 1.0, 2.0), (9, 9, 9), True)
    uo, it, _ = o.richardson_lucy(parts, psf, 3)
    ur, _, _ = r.richardson_lucy(parts, psf, 3)
    assert it == 3 and np.array_equal(uo.view(np.uint32), ur.view(np.uint32))
    # resume continues from the given estimate (APRNumerics.hpp:559-562)
    uo2, _, _ = o.richardson_lucy(parts, psf, 2, u=uo)
    ur2, _, _ = r.richardson_lucy(parts, psf, 2, u=ur)
    assert np.array_equal(uo2.view(np.uint32), ur2.view(np.uint32))


def test_richardson_lucy_constant():
    img, o = make_apr((32, 64, 64), 20, 0)
    psf = O.gaussian_stencil((1.0, 1.0, 2.0), (9, 9, 9), True)
    u, it, rel = o.richardson_lucy(np.full(o.total, 5.0, np.float32), psf, 4)
    np.testing.assert_allclose(u, 5.0, rtol=2e-5)
    u, it, rel = o.richardson_lucy(o.sample(img), psf, 200, tol=1e-3)
    assert it < 200 and rel < 1e-3 and (u > 0).all()


@pytest.mark.parametrize("shape,nb,seed", CASES[:2])
@pytest.mark.parametrize("delta", [0, -1, -2])
def test_reconstruct_matches_numpy_restatement(shape, nb, seed, delta):
    """oracle reconstruct_constant (APRReconstruction.hpp:80-120, :297-394) against the
    independent numpy piecewise-constant restatement used by the convolution ground truth"""
    img, o = make_apr(shape, nb, seed)
    parts = o.sample(img)
    t = o.tree()
    tree = o.fill_tree_mean(parts)
    level = o.l_max + delta
    ref = recon_level(o, parts.astype(np.float64), tree.astype(np.float64), t, level)
    assert not np.isnan(ref).any()
    got = o.reconstruct(parts, delta, np.float32)
    assert got.shape == ref.shape
    assert np.array_equal(got.astype(np.float64), ref.astype(np.float32).astype(np.float64))
    # u16 image: particles exact, tree means truncated like a float -> uint16 assignment
    got16 = o.reconstruct(parts, delta)
    assert np.array_equal(got16, ref.astype(np.float32).astype(np.int32).astype(np.uint16))


def test_reconstruct_then_sample_is_identity():
    """sample_image of the level-l_max reconstruction gives the particles back (every
    particle's pyramid block is constant)"""
    img, o = make_apr((29, 45, 61), 12, 3)
    parts = o.sample(img)
    rec = o.reconstruct(parts)
    assert rec.shape == img.shape
    assert np.array_equal(o.sample(rec), parts)


@pytest.mark.parametrize("shape,nb,seed", CASES[:2])
@pytest.mark.parametrize("size", [(3, 3, 3), (5, 3, 1)])
def test_local_std_equal_reference(shape, nb, seed, size):
    """APRNumerics::local_std: the oracle's restatement against the reference build; a constant
    gives a deviation at the float rounding level of E[x^2] - E[x]^2 only (the method subtracts
    two box means, so it is not exactly zero)"""
    img, o = make_apr(shape, nb, seed, "oracle")
    parts = o.sample(img)
    got = o.local_std(parts, size)
    c = o.local_std(np.full(o.total, 321, np.uint16), size)
    assert float(c.max()) < 2e-3 * 321
    pf = parts.astype(np.float32)
    assert float(got.max()) > 1.0
    if O.load("ref") is None:
        pytest.skip("reference build (oracle/_ref) not available")
    _, r = make_apr(shape, nb, seed, "ref")
    assert np.array_equal(got.view(np.uint32), r.local_std(parts, size).view(np.uint32))
    assert np.array_equal(o.local_std(pf, size).view(np.uint32), r.local_std(pf, size).view(np.uint32))


@pytest.mark.parametrize("shape,nb,seed", CASES)
def test_richardson_lucy_tv_equal_reference(shape, nb, seed):
    """APRNumerics::richardson_lucy_tv: the oracle's restatement (div_norm_grad over the oracle's
    convolution, the TV update) against the reference build, bit for bit, u16 and f32 input,
    fresh and resumed; reg = 0 is plain Richardson-Lucy"""
    img, o = make_apr(shape, nb, seed, "oracle")
    parts = o.sample(img)
    psf = O.gaussian_stencil((1.0, 1.0, 1.5), (5, 5, 5), True)
    got = o.richardson_lucy_tv(parts, psf, 3, 0.01)
    assert np.isfinite(got).all() and float(got.max()) > 1.0
    plain, _, _ = o.richardson_lucy(parts, psf, 3)
    assert np.array_equal(o.richardson_lucy_tv(parts, psf, 3, 0.0).view(np.uint32), plain.view(np.uint32))
    assert not np.array_equal(got, plain)
    if O.load("ref") is None:
        pytest.skip("reference build (oracle/_ref) not available")
    _, r = make_apr(shape, nb, seed, "ref")
    assert np.array_equal(got.view(np.uint32), r.richardson_lucy_tv(parts, psf, 3, 0.01).view(np.uint32))
    pf = parts.astype(np.float32) * np.float32(0.7)
    a = o.richardson_lucy_tv(pf, psf, 2, 0.02, u=got)
    b = r.richardson_lucy_tv(pf, psf, 2, 0.02, u=got)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _gradient_stack(o, dim, delta, sobel):
    """gradient_cfd / gradient_sobel's stencils (APRNumerics.hpp:165-213, APRStencil.hpp:135-152,
    :254-297) in float32 [z, x, y], one per level delta, rescaled by 2^-d"""
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


@pytest.mark.parametrize("shape,nb,seed", CASES[:2] + [((1, 70, 90), 6, 5)])
@pytest.mark.parametrize("sobel", [False, True])
def test_gradients_equal_reference(shape, nb, seed, sobel):
    """APRNumerics::gradient_cfd / gradient_sobel / gradient_magnitude_*: the stencils restated
    here (incl. the 2-D Sobel of a one-plane image) with the oracle's convolution, against the
    reference build's own functions"""
    need_ref()
    img, o = make_apr(shape, nb, seed, "oracle")
    r = O.apr_from_structure(o.dims, o.xz_end_vec, o.y_vec, "ref")
    if not hasattr(r.lib, "orc_gradient_u16"):
        pytest.skip("reference build without the gradient exports")
    parts = o.sample(img)
    pf = (parts.astype(np.float32) * np.float32(0.731)).astype(np.float32)
    dims = [d for d in (0, 1, 2) if o.dims[d] > 1]
    grads = {}
    for dim, delta in ((0, 1.0), (1, 0.7), (2, 2.5)):
        if dim not in dims:
            continue
        stack = _gradient_stack(o, dim, delta, sobel)
        for p in (parts, pf):
            exp = o.convolve_pencil(stack, p, True)
            assert np.array_equal(exp.view(np.uint32), r.gradient(p, sobel, dim, delta).view(np.uint32)), (dim, p.dtype)
        grads[dim] = o.convolve_pencil(stack, pf, True)
    m = grads[0] * grads[0]
    for dim in dims[1:]:
        m = (m + grads[dim] * grads[dim]).astype(np.float32)
    m = np.sqrt(m).astype(np.float32)
    assert np.array_equal(m.view(np.uint32), r.gradient_magnitude(pf, sobel, (1.0, 0.7, 2.5)).view(np.uint32))


FACE_CASES = CASES + [((1, 70, 90), 6, 5), ((40, 1, 50), 5, 2), ((33, 17, 9), 5, 6), ((61, 63, 65), 30, 1)]


@pytest.mark.parametrize("shape,nb,seed", FACE_CASES)
def test_face_neighbour_filters_equal_reference(shape, nb, seed):
    """APRNumerics::face_neighbour_filter per dimension and seperable_face_neighbour_filter
    (0, 1, 2 repeats): the oracle's restatement of the random iterator's neighbour search
    against the reference build, bit for bit, on odd, even and one-plane images"""
    need_ref()
    img, o = make_apr(shape, nb, seed, "oracle")
    r = O.apr_from_structure(o.dims, o.xz_end_vec, o.y_vec, "ref")
    if not hasattr(r.lib, "orc_face_neighbour_filter_f32"):
        pytest.skip("reference build without the face-neighbour export")
    pf = (o.sample(img).astype(np.float32) * np.float32(0.37)).astype(np.float32)
    for f in (np.array([0.25, 0.5, 0.3], np.float32), np.array([1.0, 0.0, 0.0], np.float32)):
        for dim in range(3):
            a, b = o.face_neighbour_filter(pf, f, dim), r.face_neighbour_filter(pf, f, dim)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (dim, int((a != b).sum()))
        for rep in (0, 1, 2):
            a, b = o.face_neighbour_filter(pf, f, 0, rep), r.face_neighbour_filter(pf, f, 0, rep)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (rep, int((a != b).sum()))


def test_face_neighbour_filter_closed_forms():
    """a constant stays constant under weights summing to 1; the identity filter is the input"""
    img, o = make_apr((29, 45, 61), 12, 3, "oracle")
    c = np.full(o.total, 3.5, np.float32)
    f = np.array([0.25, 0.5, 0.25], np.float32)
    for dim in range(3):
        assert np.all(o.face_neighbour_filter(c, f, dim) == np.float32(3.5))
    p = o.sample(img).astype(np.float32)
    assert np.array_equal(o.face_neighbour_filter(p, np.array([0, 1, 0], np.float32), 1), p)


def _sat_windows_f64(img, lev, scale, om_in, d_max, axis):
    """One adaptive mean pass of reconstruct_smooth written as explicit windows in float64 (the
    mean of each pixel's window, not the reference's cumulative sums): half-width
    off = min(floor(2^(d_max - level) / scale), om), om = min(om_in, n // 2 - 1); along y the
    window is [0, off] at k = 0, [1, k + off] for 1 <= k <= off, [k - off, k + off] inside and
    [k - off, n - 1] at the high end; along x / z it is [0, i + off] at the low end (i <= off;
    along z only where the x index is <= off, elsewhere the pixel keeps its value),
    [i - off, i + off] inside and [i - off, n - 1] at the high end (APRReconstruction.hpp:528-876)"""
    a = np.moveaxis(img.astype(np.float64), 2 - axis, -1)          # [.., n] along the pass axis
    L = np.moveaxis(lev, 2 - axis, -1)
    n = a.shape[-1]
    om = min(om_in, n // 2 - 1)
    out = a.copy()
    offs = np.minimum(np.floor(2.0 ** (d_max - L.astype(np.int64)) / np.float64(np.float32(scale))), om).astype(np.int64)
    xidx = np.moveaxis(np.broadcast_to(np.arange(img.shape[1])[None, :, None], img.shape), 2 - axis, -1)
    for idx in np.ndindex(a.shape[:-1]):
        line, o = a[idx], offs[idx]
        for i in range(n):
            f = int(o[i])
            if axis == 0:
                if i == 0:
                    lo, hi = 0, f
                elif i <= f:
                    lo, hi = 1, i + f
                elif i < n - f:
                    lo, hi = i - f, i + f
                else:
                    lo, hi = i - f, n - 1
            else:
                if i >= n - f:
                    lo, hi = i - f, n - 1
                elif i >= f + 1:
                    lo, hi = i - f, i + f
                elif axis == 1 or xidx[idx][i] < f + 1:
                    lo, hi = 0, i + f
                else:
                    continue
            out[idx + (i,)] = line[lo:hi + 1].mean()
    return np.moveaxis(out, -1, 2 - axis)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_reconstruct_smooth_pass_windows(axis):
    """each adaptive pass of the oracle equals its window means (float64, 1e-5 relative); the
    level image mixes every half-width from 0 to the clamp"""
    rng = np.random.default_rng(10 + axis)
    shape = (30, 34, 38)                                   # z, x, y: every axis > 2 * om + 2
    img = rng.integers(0, 100, shape).astype(np.float32)      # integer samples: the float prefix sums are exact
    lev = rng.integers(2, 7, shape).astype(np.uint16)     # d_max 6: 2^(6 - l) / 2 = 0.5 .. 8
    got = img.copy()
    O.load("oracle").orc_sat_adaptive_f32(got.reshape(-1), lev.reshape(-1), shape[2], shape[1], shape[0], 2.0, 20, 6, axis)
    exp = _sat_windows_f64(img, lev, 2.0, 20, 6, axis)
    err = np.abs(got - exp) / np.maximum(np.abs(exp), 1.0)
    assert err.max() < 1e-5, (axis, float(err.max()))
    # the clamp: offset_max 3 caps every half-width
    got = img.copy()
    O.load("oracle").orc_sat_adaptive_f32(got.reshape(-1), lev.reshape(-1), shape[2], shape[1], shape[0], 0.25, 3, 6, axis)
    exp = _sat_windows_f64(img, lev, 0.25, 3, 6, axis)
    assert (np.abs(got - exp) / np.maximum(np.abs(exp), 1.0)).max() < 1e-5


def test_reconstruct_smooth_closed_forms():
    """half-width 0 everywhere (scale far above 2^(l_max - l_min)) leaves the constant
    reconstruction unchanged; a constant stays constant (float) for any scale"""
    img, o = make_apr((29, 45, 61), 12, 3, "oracle")
    p = o.sample(img)
    assert np.array_equal(o.reconstruct_smooth(p, scale=(1e6, 1e6, 1e6)), o.reconstruct(p))
    c = np.full(o.total, 7.25, np.float32)
    for sc in ((2.0, 2.0, 2.0), (1.0, 0.5, 3.0)):
        out = o.reconstruct_smooth(c, scale=sc)
        assert np.abs(out - np.float32(7.25)).max() < 1e-5 * 7.25, sc
