"""Double-precision restatements of the float stages, to state the error of each float result
(test infrastructure, CPU only).

These compute what the float32 code computes, in float64 and in closed form where the
reference's float code accumulates (window means as sums over the window, not running
prefix sums):

* bspline_smooth: the smoothing spline of ComputeGradient.hpp:248-779 (Unser's recursion,
  coefficients :256-330) in float64, float input, no rounding (APRConverter<float>);
* gradient_ds: |grad| by central differences with replicated ends, max over 2x2x2
  (ComputeGradient.hpp:975-1051);
* mean_ds: 2x2x2 mean (PixelData.hpp:915-967);
* inv_bspline: [1, 4, 1] / 6 per axis with the reference's end rows (:784-964);
* local_intensity_scale: reflect padding, truncated-window means, |I - mean|, second mean,
  rescale (LocalIntensityScale.hpp:26-494);
* tree_mean: each tree node the mean of its children (particles or tree nodes one level
  finer), the closed form of APRTreeNumerics::fill_tree_mean's boundary scale factors
  (APRTreeNumerics.hpp:63-202).
"""
import math

import numpy as np


def bspline_coeffs(lam=3.0, tol=1e-4, n=None):
    """b1, b2, norm, bc1..bc4 (ComputeGradient.hpp:256-330) in float64"""
    xi = 1 - 96 * lam + 24 * lam * math.sqrt(3 + 144 * lam)
    rho = (24 * lam - 1 - math.sqrt(xi)) / (24 * lam) * math.sqrt((1 / xi) * (48 * lam + 24 * lam * math.sqrt(3 + 144 * lam)))
    omg = math.atan(math.sqrt((1 / xi) * (144 * lam - 1)))
    r2 = rho * rho
    c0 = (1 + r2) / (1 - r2) * (1 - 2 * rho * math.cos(omg) + r2) / (1 + 2 * rho * math.cos(omg) + r2)
    gamma = (1 - r2) / (1 + r2) / math.tan(omg)
    b1, b2 = 2 * rho * math.cos(omg), -r2
    k0 = int(math.ceil(abs(math.log(tol) / math.log(rho))))
    norm = (1 - 2 * rho * math.cos(omg) + r2) ** 2
    k = np.arange(k0 + 3, dtype=np.float64)
    f = rho ** k * np.sin((k + 1) * omg) / math.sin(omg)
    b = c0 * rho ** k * (np.cos(omg * k) + gamma * np.sin(omg * k)) / (1 - 2 * rho * math.cos(omg) + r2) ** 2
    mL = min(k0, n) if n is not None else k0
    bc1 = np.zeros(k0)
    bc1[1] = f[0]
    bc1 += f[1:k0 + 1]
    bc2 = f[:k0].copy()
    bc3 = np.zeros(k0)
    bc3[0] = b[1]
    bc3[1:] += b[:k0 - 1] + b[2:k0 + 1]
    bc4 = 2.0 * b[:k0]
    bc4[0] = b[0]
    out = []
    for bc in (bc1, bc2, bc3, bc4):
        bc = bc.copy()
        bc[mL - 1] += bc[mL:].sum()
        out.append(bc[:mL])
    return dict(b1=b1, b2=b2, norm=norm, k0=k0, rho=rho, omg=omg, bc=out)


def _bspline_axis(a, axis, lam, tol):
    x = np.moveaxis(a, axis, -1).copy()
    n = x.shape[-1]
    c = bspline_coeffs(lam, tol, n)
    bc1, bc2, bc3, bc4 = c["bc"]
    b1, b2, nf = c["b1"], c["b2"], c["norm"]
    mL = len(bc1)
    t1 = x[..., :mL] @ bc1
    t2 = x[..., :mL] @ bc2
    rev = x[..., ::-1]
    t3 = rev[..., :mL] @ bc3
    t4 = rev[..., :mL] @ bc4
    out = np.empty_like(x)
    out[..., 0], out[..., 1] = t2, t1
    for k in range(2, n):
        t = t1 * b1 + t2 * b2 + x[..., k]
        out[..., k] = t
        t2, t1 = t1, t
    out[..., n - 2] = t3 * nf
    out[..., n - 1] = t4 * nf
    s2, s1 = out[..., n - 1] / nf, out[..., n - 2] / nf
    for k in range(n - 3, -1, -1):
        t = s1 * b1 + s2 * b2 + out[..., k]
        out[..., k] = t * nf
        s2, s1 = s1, t
    return np.moveaxis(out, -1, axis)


def bspline_smooth(img_zxy, lam=3.0, tol=1e-4):
    """get_smooth_bspline_3D on a float image [z, x, y]: y, then x, then z (axes longer than 2)"""
    a = img_zxy.astype(np.float64)
    for axis in (2, 1, 0):
        if a.shape[axis] > 2:
            a = _bspline_axis(a, axis, lam, tol)
    return a


def _nb(a, axis, d):
    """a shifted by d along axis with replicated ends"""
    n = a.shape[axis]
    idx = np.clip(np.arange(n) + d, 0, n - 1)
    return np.take(a, idx, axis=axis)


def _ds_index(n):
    """the two fine indices of each coarse cell: 2k and min(2k + 1, n - 1)"""
    k = np.arange((n + 1) // 2)
    return 2 * k, np.minimum(2 * k + 1, n - 1)


def _ds(a, op):
    iz, ix, iy = (_ds_index(n) for n in a.shape)
    vals = [a[z][:, x][:, :, y] for z in iz for x in ix for y in iy]
    return op(vals)


def gradient_ds(s_zxy, h=(1.0, 1.0, 1.0)):
    """|grad| (central differences over 2 h, replicated ends), then the 2x2x2 maximum; h = (hy, hx, hz)"""
    hy, hx, hz = h
    gy = (_nb(s_zxy, 2, 1) - _nb(s_zxy, 2, -1)) / (2 * hy)
    gx = (_nb(s_zxy, 1, 1) - _nb(s_zxy, 1, -1)) / (2 * hx)
    gz = (_nb(s_zxy, 0, 1) - _nb(s_zxy, 0, -1)) / (2 * hz)
    g = np.sqrt(gx * gx + gz * gz + gy * gy)
    return _ds(g, lambda v: np.maximum.reduce(v))


def mean_ds(s_zxy):
    return _ds(s_zxy, lambda v: sum(v) / 8.0)


def inv_bspline(a_zxy):
    a = a_zxy.astype(np.float64)
    for axis in (2, 1, 0):
        n = a.shape[axis]
        if n <= 2:
            continue
        x = np.moveaxis(a, axis, -1)
        o = np.empty_like(x)
        o[..., 0] = 4 / 6 * x[..., 0] + 2 / 6 * x[..., 1]
        o[..., 1:-1] = (x[..., :-2] + 4 * x[..., 1:-1] + x[..., 2:]) / 6
        o[..., -1] = 2 / 6 * x[..., -2] + 4 / 6 * x[..., -1]
        a = np.moveaxis(o, -1, axis)
    return a


def _window_mean(a, axis, off):
    """mean over [i - off, i + off] clipped to the axis (the running-sum means' ends)"""
    x = np.moveaxis(a, axis, -1)
    n = x.shape[-1]
    c = np.concatenate([np.zeros(x.shape[:-1] + (1,)), np.cumsum(x, axis=-1)], axis=-1)
    i = np.arange(n)
    lo, hi = np.maximum(i - off, 0), np.minimum(i + off, n - 1)
    o = (c[..., hi + 1] - c[..., lo]) / (hi - lo + 1)
    return np.moveaxis(o, -1, axis)


def local_intensity_scale(i_zxy, psf_ind=1):
    """sigma = rescale * mean2(|I - mean1(I)|) on the reflect-padded array (the PSF of 2 px:
    windows 1 and 3, 3D rescale 26.1256; every axis longer than 3)"""
    w1 = [1, 1, 1, 2, 2, 3][psf_ind]
    w2 = [2, 3, 4, 4, 5, 6][psf_ind]
    r3 = [12.8214, 26.1256, 40.2795, 23.3692, 36.2061, 27.0385][psf_ind]
    wv = max(w1, w2)
    assert all(n > wv for n in i_zxy.shape)
    pads = [min(wv, n - 1) for n in i_zxy.shape]
    a = np.pad(i_zxy.astype(np.float64), [(p, p) for p in pads], mode="reflect")
    b = a.copy()
    for axis in (2, 1, 0):
        a = _window_mean(a, axis, w1)
    a = np.abs(a - b)
    for axis in (2, 1, 0):
        a = _window_mean(a, axis, w2)
    a = a * np.float64(np.float32(r3))
    pz, px, py = pads
    return a[pz:a.shape[0] - pz, px:a.shape[1] - px, py:a.shape[2] - py]


def tree_mean(apr_dims, l_min, l_max, part_coords, parts, tree_coords):
    """tree node values in tree order: the mean of each node's children, finest tree level
    first.  part_coords / tree_coords: (level, z, x, y) arrays."""
    y, x, z = apr_dims
    plv, pz, px, py = part_coords
    tlv, tz, tx, ty = tree_coords
    vals = np.full(len(tlv), np.nan)
    parts = parts.astype(np.float64)
    for t in range(int(tlv.max()) if len(tlv) else -1, int(tlv.min()) - 1 if len(tlv) else -1, -1):
        c = t + 1
        s = 2 ** (l_max - c)
        dims = (-(-z // s), -(-x // s), -(-y // s))
        D = np.full(dims, np.nan)
        m = plv == c
        D[pz[m], px[m], py[m]] = parts[m]
        m = tlv == c
        D[tz[m], tx[m], ty[m]] = vals[m]
        m = np.nonzero(tlv == t)[0]
        for i in m:
            blk = D[2 * tz[i]:2 * tz[i] + 2, 2 * tx[i]:2 * tx[i] + 2, 2 * ty[i]:2 * ty[i] + 2]
            vals[i] = np.nanmean(blk)
    return vals
