"""GPU parity of the blocked conversion (APRConverterBatch, get_apr over z blocks with ghost
slabs) and of the blocked sampling against the CPU oracle, through the C-ABI: structure and
particle intensities bit-exact, for block and ghost sizes that do and do not align with the
coarse levels; with ghosts covering the whole image the result is plain get_apr's (the
reference's own check, APRTest.cpp:2869-2928)."""
import numpy as np
import pytest

import libapr_amd as A

pytestmark = pytest.mark.gpu

# (z, x, y), blobs, seed, z_block_size, ghost_z
CASES = [
    ((64, 72, 80), 30, 3, 16, 64),     # ghosts cover the image: plain get_apr
    ((64, 72, 80), 30, 3, 16, 8),
    ((61, 40, 50), 30, 3, 13, 5),      # uneven blocks, ghosts not aligned with any level
    ((37, 33, 45), 30, 3, 8, 0),       # no ghosts
    ((70, 50, 66), 40, 4, 9, 3),       # blocks whose coarse rows reach past a level's end
    ((20, 30, 40), 10, 5, 100, 4),     # one block
    ((96, 40, 128), 60, 6, 32, 16),
]


def _par(O):
    return O.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)


def _apar():
    return A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)


@pytest.mark.parametrize("shape,nb,seed,zbs,ghost", CASES)
def test_get_apr_by_block_matches_oracle(oracle, ctx, shape, nb, seed, zbs, ghost):
    img = A.synth_blobs(shape, seed, nb)
    ref = oracle.get_apr_by_block(img, _par(oracle), zbs, ghost)
    conv = A.APRConverterBatch(ctx)
    conv.par = _apar()
    conv.z_block_size, conv.ghost_z = zbs, ghost
    apr = conv.get_apr(img)
    print(f"case {shape} blocks {zbs} ghost {ghost}: P={apr.total} oracle={ref.total}", flush=True)
    assert apr.total == ref.total
    lxz, xze, yv = apr.structure()
    assert np.array_equal(lxz, ref.level_xz_vec)
    assert np.array_equal(xze, ref.xz_end_vec)
    assert np.array_equal(yv, ref.y_vec)
    parts = A.ParticleData().sample_parts_from_img_blocked(apr, img, zbs, ghost)
    assert np.array_equal(parts, ref.sample_blocked(img, zbs, ghost))
    if ghost >= shape[0]:
        plain = A.APRConverter(ctx)
        plain.par = _apar()
        p = plain.get_apr(img)
        assert p.total == apr.total and np.array_equal(p.structure()[1], xze) and np.array_equal(p.structure()[2], yv)
        assert np.array_equal(parts, A.ParticleData().sample_image(p, img))
        p.close()
    apr.close()


def test_get_apr_by_block_device_image(oracle, ctx):
    shape, zbs, ghost = (48, 64, 96), 12, 6
    img = A.synth_blobs(shape, 9, 30)
    ref = oracle.get_apr_by_block(img, _par(oracle), zbs, ghost)
    d = A.DeviceArray(img.nbytes, ctx)
    d.from_host(img)
    conv = A.APRConverterBatch(ctx)
    conv.par = _apar()
    conv.z_block_size, conv.ghost_z = zbs, ghost
    apr = conv.get_apr(d.ptr, dims=shape[::-1])
    assert apr.total == ref.total and np.array_equal(apr.structure()[2], ref.y_vec)
    out = A.DeviceArray(apr.total * 2, ctx)
    A.ParticleData().sample_parts_from_img_blocked(apr, d.ptr, zbs, ghost, out=out.ptr)
    assert np.array_equal(out.to_host(np.uint16, apr.total), ref.sample_blocked(img, zbs, ghost))
    apr.close()


def test_get_apr_by_block_rejects_bad_sizes(ctx):
    conv = A.APRConverterBatch(ctx)
    conv.z_block_size = 0
    with pytest.raises(A.AprhError) as e:
        conv.get_apr(A.synth_blobs((8, 16, 16), 1, 2))
    assert e.value.code == A.APRH_ERR_INVALID_ARGUMENT
