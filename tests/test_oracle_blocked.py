"""The oracle's restatement of the blocked conversion and blocked sampling
(APRConverterBatch.hpp:111-175, PullingScheme.hpp:175-244, ParticleData.hpp:292-362)
pinned two ways: against the reference's own APRConverterBatch and
sample_parts_from_img_blocked compiled from the reference tree (oracle/_ref/libref_batch.so,
which reads its input from a TIFF file), and by the reference test's property that ghosts
covering the whole image reproduce plain get_apr and sample_image (APRTest.cpp:2869-2928).

The reference build is compared on block / ghost sizes whose coarse rows stay inside each
level: for others its blocked sampler and fill_patch step past a level's rows (into the next
level's particles, or past the particle-cell tree's planes), which the restatement skips."""
import ctypes as C
import os

import numpy as np
import pytest

import libapr_amd as A
import oracle_lib as O

REF_BATCH = os.path.join(O.ORACLE_DIR, "_ref", "libref_batch.so")

CASES = [((64, 72, 80), 16, 64), ((61, 40, 50), 13, 5), ((37, 33, 45), 8, 0), ((20, 30, 40), 100, 4),
         ((96, 40, 128), 24, 8), ((48, 64, 96), 12, 4), ((128, 32, 64), 12, 8), ((80, 48, 60), 32, 16)]


def _par():
    return O.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)


def _l_max(shape):
    return max(1, int(np.ceil(np.log2(max(shape)))))


def stays_inside(shape, zbs, ghost):
    """every block's rows of every level lie inside the level (no row past its z extent)"""
    z = shape[0]
    lm = _l_max(shape)
    nb = max(z // zbs, 1)
    for b in range(nb):
        z0 = b * zbs
        zf = z if b == nb - 1 else (b + 1) * zbs
        gl, gr = min(z0, ghost), min(z - zf, ghost)
        zp = zf + gr - (z0 - gl)
        for lev in range(1, lm + 1):
            f = 2 ** (lm - lev)
            zn, zl, zr = -(-zp // f), gl // f, gr // f
            if zn - zr + z0 // f - zl > -(-z // f):
                return False
    return True


def _ref_batch():
    if not os.path.exists(REF_BATCH):
        return None
    L = C.CDLL(REF_BATCH)
    u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
    L.refb_get_apr_by_block_u16.argtypes = [u16p, C.c_int64, C.c_int64, C.c_int64, C.POINTER(O.AprhParams), C.c_int64,
                                            C.c_int64]
    L.refb_get_apr_by_block_u16.restype = C.c_void_p
    L.refb_info.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]
    L.refb_get.argtypes = [C.c_void_p, u64p, u64p, u16p]
    L.refb_sample_blocked.argtypes = [C.c_void_p, C.c_int64, C.c_int64, u16p]
    L.refb_free.argtypes = [C.c_void_p]
    return L


@pytest.mark.parametrize("shape,zbs,ghost", CASES)
def test_blocked_restatement_equals_reference_build(oracle, shape, zbs, ghost):
    L = _ref_batch()
    if L is None:
        pytest.skip("oracle/_ref/libref_batch.so not built (reference tree or libtiff absent)")
    assert stays_inside(shape, zbs, ghost)
    img = A.synth_blobs(shape, 3, 30)
    z, x, y = shape
    par = _par()
    h = L.refb_get_apr_by_block_u16(img.reshape(-1), y, x, z, C.byref(par), zbs, ghost)
    assert h
    try:
        info = np.zeros(3, np.int64)
        L.refb_info(h, info)
        lxz, xze = np.zeros(info[2], np.uint64), np.zeros(info[1], np.uint64)
        yv = np.zeros(max(info[0], 1), np.uint16)
        L.refb_get(h, lxz, xze, yv)
        ps = np.zeros(max(info[0], 1), np.uint16)
        L.refb_sample_blocked(h, zbs, ghost, ps)
    finally:
        L.refb_free(h)
    o = oracle.get_apr_by_block(img, par, zbs, ghost)
    assert o.total == info[0]
    assert np.array_equal(o.level_xz_vec, lxz) and np.array_equal(o.xz_end_vec, xze)
    assert np.array_equal(o.y_vec, yv[:o.total])
    assert np.array_equal(o.sample_blocked(img, zbs, ghost), ps[:o.total])


@pytest.mark.parametrize("shape,zbs", [((64, 72, 80), 16), ((45, 30, 70), 11)])
def test_full_ghosts_give_plain_get_apr(oracle, shape, zbs):
    img = A.synth_blobs(shape, 7, 25)
    par = _par()
    plain = oracle.get_apr(img, par)
    blk = oracle.get_apr_by_block(img, par, zbs, shape[0])
    assert blk.total == plain.total
    assert np.array_equal(blk.xz_end_vec, plain.xz_end_vec) and np.array_equal(blk.y_vec, plain.y_vec)
    assert np.array_equal(blk.sample_blocked(img, zbs, shape[0]), plain.sample(img))


def test_one_block_gives_plain_get_apr(oracle):
    img = A.synth_blobs((20, 30, 40), 5, 10)
    par = _par()
    plain = oracle.get_apr(img, par)
    blk = oracle.get_apr_by_block(img, par, 64, 0)
    assert np.array_equal(blk.y_vec, plain.y_vec) and np.array_equal(blk.sample_blocked(img, 64, 0), plain.sample(img))


def test_overrunning_blocks_are_detected():
    # the GPU test's case with coarse rows past a level's end (the reference build is not run on it)
    assert not stays_inside((70, 50, 66), 9, 3)
    assert not stays_inside((64, 72, 80), 16, 8)
