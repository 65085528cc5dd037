"""aprh_sample_image_u16_noalloc (sample_image with its mean pyramid in a fixed device array)
against aprh_sample_image_u16, bit for bit, through the C-ABI: two calls in a row on one
context, the second on a smaller volume, so the second call runs over the first call's
scratch contents.  Inputs and outputs are device memory, as that entry point requires."""
import ctypes as C

import numpy as np
import pytest

import libapr_amd as A

pytestmark = pytest.mark.gpu


def _sample_both(ctx, shape, seed, nb):
    z, x, y = shape
    L = A.lib()
    img = A.DeviceArray(z * x * y * 2, ctx)
    A.synth_blobs(shape, seed, nb, out=img.ptr, ctx=ctx)
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)
    apr = conv.get_apr(img.ptr, dims=(y, x, z))
    P = apr.total
    ref = A.DeviceArray(max(P, 1) * 2, ctx)
    got = A.DeviceArray(max(P, 1) * 2, ctx)
    got.from_host(np.full(max(P, 1), 0xBEEF, np.uint16))   # anything the call fails to write shows up
    A._check(L.aprh_sample_image_u16(ctx.h, apr.h, C.c_void_p(img.ptr), C.c_void_p(ref.ptr)))
    A._check(L.aprh_sample_image_u16_noalloc(ctx.h, apr.h, C.c_void_p(img.ptr), C.c_void_p(got.ptr)))
    r, g = ref.to_host(np.uint16, P), got.to_host(np.uint16, P)
    print(f"noalloc case {shape} P={P}", flush=True)
    apr.close()
    for a in (img, ref, got):
        a.free()
    return r, g


def test_noalloc_sample_matches_sample_image_twice():
    ctx = A.Context(0)
    try:
        assert A.lib().aprh_noalloc_scratch_bytes() == 64 << 20
        # first a larger volume, then a smaller one: the second call reads nothing stale
        for shape, seed, nb in (((96, 160, 192), 11, 120), ((24, 40, 72), 12, 10)):
            r, g = _sample_both(ctx, shape, seed, nb)
            assert r.size > 0
            assert np.array_equal(r, g), (shape, int((r != g).sum()))
    finally:
        ctx.close()


def test_noalloc_rejects_host_buffers():
    ctx = A.Context(0)
    try:
        shape = (8, 16, 32)
        img = A.synth_blobs(shape, 3, 2)
        conv = A.APRConverter(ctx)
        conv.par = A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)
        apr = conv.get_apr(img)
        out = np.zeros(max(apr.total, 1), np.uint16)
        rc = A.lib().aprh_sample_image_u16_noalloc(ctx.h, apr.h, C.c_void_p(img.ctypes.data), C.c_void_p(out.ctypes.data))
        assert rc == A.APRH_ERR_INVALID_ARGUMENT
        apr.close()
    finally:
        ctx.close()
