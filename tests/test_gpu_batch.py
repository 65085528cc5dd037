"""GPU parity of the batched conversion (one HIP graph over many volumes of
uneven sizes) against single-volume get_apr and the CPU oracle: particle counts,
level_xz_vec, xz_end_vec and y_vec bit-exact per volume, over two replays of
the same graph with different images."""
import ctypes as C

import numpy as np
import pytest

import libapr_amd as A

pytestmark = pytest.mark.gpu

# (y, x, z): odd and even sides, flat and tall volumes, l_max from 6 to 8
DIMS = [(64, 48, 32), (40, 72, 24), (129, 33, 17), (96, 96, 64), (20, 200, 9), (33, 31, 35)]


def _par(O):
    return O.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)


def _apar():
    return A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)


def _images(seed0):
    out = []
    for i, (y, x, z) in enumerate(DIMS):
        nb = max(2, y * x * z // 32 ** 3 * 4)
        out.append(A.synth_blobs((z, x, y), seed0 + i, nb))
    return out


def test_batch_matches_single_and_oracle(oracle, ctx):
    batch = A.APRBatch(DIMS, par=_apar(), ctx=ctx)
    stages = batch.stages()
    # one launch per stage over the volumes: far fewer launches than volumes x stages
    assert 0 < len(stages) < 3 * len(DIMS) * 10
    assert all(1 <= st["grid"] <= st["tiles"] and 1 <= st["volumes"] <= len(DIMS) for st in stages)
    assert max(st["volumes"] for st in stages) == len(DIMS)
    conv = A.APRConverter(ctx)
    conv.par = _apar()
    for rep, seed0 in enumerate((10, 50)):
        imgs = _images(seed0)
        for i, img in enumerate(imgs):
            batch.set_input(i, img)
        print(f"batch replay {rep}: {len(DIMS)} volumes", flush=True)
        batch.run()
        for i, img in enumerate(imgs):
            got = batch.get(i)
            single = conv.get_apr(img)
            ref = oracle.get_apr(img, _par(oracle), "oracle")
            print(f"  volume {i} dims {DIMS[i]}: P={got.total} single={single.total} oracle={ref.total}", flush=True)
            assert got.total == single.total == ref.total
            assert got.bspline_offset == single.bspline_offset
            lxz, xze, yv = got.structure()
            assert np.array_equal(lxz, ref.level_xz_vec)
            assert np.array_equal(xze, ref.xz_end_vec)
            assert np.array_equal(yv, ref.y_vec)
            parts = A.ParticleData().sample_image(got, img)
            assert np.array_equal(parts, ref.sample(img))
            got.close()
            single.close()
    batch.close()


def test_batch_mixes_long_and_short_rows(oracle, ctx):
    # a volume with rows of 1024+ cells at the top level (64-lane row fill) merged into one plan
    # with small volumes (16-lane row fill)
    dims = [(2048, 8, 4), (64, 48, 32), (1030, 10, 6), (40, 72, 24)]
    batch = A.APRBatch(dims, par=_apar(), ctx=ctx)
    imgs = []
    for i, (y, x, z) in enumerate(dims):
        imgs.append(A.synth_blobs((z, x, y), 70 + i, 6))
        batch.set_input(i, imgs[-1])
    batch.run()
    for i, img in enumerate(imgs):
        got = batch.get(i)
        ref = oracle.get_apr(img, _par(oracle), "oracle")
        print(f"volume {i} dims {dims[i]}: P={got.total} oracle={ref.total}", flush=True)
        assert got.total == ref.total
        lxz, xze, yv = got.structure()
        assert np.array_equal(lxz, ref.level_xz_vec)
        assert np.array_equal(xze, ref.xz_end_vec)
        assert np.array_equal(yv, ref.y_vec)
        got.close()
    batch.close()


def test_batch_rejects_small_volumes(ctx):
    with pytest.raises(A.AprhError) as e:
        A.APRBatch([(64, 64, 64), (4, 4, 4)], ctx=ctx)
    assert e.value.code == A.APRH_ERR_UNSUPPORTED


def test_batch_apr_is_borrowed(ctx):
    batch = A.APRBatch([(40, 40, 24)], par=_apar(), ctx=ctx)
    batch.set_input(0, A.synth_blobs((24, 40, 40), 3, 4))
    batch.run()
    apr = batch.get(0)
    assert A.lib().aprh_apr_destroy(apr.h) == A.APRH_ERR_INVALID_ARGUMENT
    assert apr.total > 0
    batch.close()


def test_batch_offset_rerun_per_volume(oracle, ctx):
    # volumes whose lines are a multiple of 64 find the bspline offset in the y pass: a guess,
    # then a second run inside the graph for a volume whose (min, max) give another offset;
    # the two replays swap which volume needs it
    dims = [(128, 40, 24), (64, 72, 20)]
    par = A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10, lambda_=0.2)
    opar = oracle.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10, lambda_=0.2)

    def image(i, mid):
        y, x, z = dims[i]
        img = A.synth_blobs((z, x, y), 20 + i, 12).astype(np.float64)
        lo, hi = (20000, 47000) if mid else (100, 5000)
        return (lo + (img - img.min()) * (hi - lo) / max(img.max() - img.min(), 1.0)).round().astype(np.uint16)

    batch = A.APRBatch(dims, par=par, ctx=ctx)
    conv = A.APRConverter(ctx)
    conv.par = par
    for rep in range(2):
        imgs = [image(0, rep == 0), image(1, rep == 1)]
        for i, img in enumerate(imgs):
            batch.set_input(i, img)
        batch.run()
        for i, img in enumerate(imgs):
            got = batch.get(i)
            single = conv.get_apr(img)
            ref = oracle.get_apr(img, opar, "oracle")
            print(f"replay {rep} volume {i}: offset {got.bspline_offset} P={got.total} oracle={ref.total}", flush=True)
            assert got.bspline_offset == single.bspline_offset
            assert got.total == single.total == ref.total
            lxz, xze, yv = got.structure()
            assert np.array_equal(xze, ref.xz_end_vec) and np.array_equal(yv, ref.y_vec)
            got.close()
            single.close()
    batch.close()
