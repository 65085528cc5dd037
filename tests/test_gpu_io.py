"""APR files with device APRs (APRFile.hpp:178-356), through the C-ABI: the GPU conversion of
the reference test's sphere image equals the structure of the reference's own sphere.apr
(tests/golden/<case>.apr, APRTest.cpp:2816-2836 makes the same comparison), and a GPU APR
written with its tree and particles reads back unchanged, also into a second context."""
import os

import numpy as np
import pytest

import libapr_amd as A

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(A.IO_LIB_PATH), reason="libaprh_io.so not built")]

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _fixture(case):
    d = np.load(os.path.join(GOLD, case + ".npz"))
    return {k: d[k] for k in d.files}


@pytest.mark.parametrize("case", ["sphere_120", "sphere_diff_dims"])
def test_read_reference_file_into_device_apr(ctx, case):
    f = _fixture(case)
    img = np.ascontiguousarray(f["original"])
    af = A.APRFile()
    af.open(os.path.join(GOLD, case + ".apr"), "READ")
    apr = af.read_apr(ctx=ctx)
    parts = af.read_particles("particle_intensities")
    af.close()
    print(f"{case}: P={apr.total}", flush=True)
    assert apr.total == int(f["attr_total_number_particles"])
    # the reference's structure, sampled on the GPU, gives the file's intensities
    assert np.array_equal(A.ParticleData().sample_image(apr, img), parts)
    # and the GPU's conversion from the test's saved gradient and scale gives the file's structure
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(Ip_th=0.0, rel_error=float(f["attr_rel_error"]), lambda_=float(f["attr_lambda"]),
                               sigma_th_max=max(float(f["attr_sigma_th_max"]), 0.1),
                               sigma_th=max(float(f["attr_sigma_th"]), 0.1), grad_th=float(f["attr_grad_th"]))
    Z, X, Y = img.shape
    g = conv.get_apr_custom_grad_scale(f["gradient_saved"], f["scale_saved"], np.ones(f["scale_saved"].shape, np.float32),
                                       (Y, X, Z))
    for a, b in zip(g.structure(), apr.structure()):
        assert np.array_equal(a, b)
    g.close()
    apr.close()


def test_write_read_device_apr(ctx, tmp_path):
    img = A.synth_blobs((40, 72, 96), 3, 24)
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)
    apr = conv.get_apr(img)
    parts = A.ParticleData().sample_image(apr, img)
    d_parts = A.DeviceArray(parts.nbytes, ctx)
    d_parts.from_host(parts)
    tree = apr.tree()
    path = str(tmp_path / "gpu.apr")
    af = A.APRFile()
    af.open(path, "WRITE")
    af.write_apr(apr, write_tree=True, parameters=conv.par)
    d_parts.size = parts.size
    af.write_particles("particles", d_parts, ctx=ctx, dtype="uint16")
    tm = A.APRTreeNumerics.fill_tree_mean(apr, parts)
    af.write_particles("tree_means", tm, apr_or_tree=False)
    af.close()
    other = A.Context(0)
    af.open(path, "READ")
    back = af.read_apr(ctx=other)
    assert back.total == apr.total and back.org_dims == apr.org_dims
    for a, b in zip(apr.structure(), back.structure()):
        assert np.array_equal(a, b)
    assert np.array_equal(af.read_particles("particles"), parts)
    assert np.array_equal(af.read_particles("tree_means", apr_or_tree=False).view(np.uint32), tm.view(np.uint32))
    af.close()
    t2 = back.tree()
    assert np.array_equal(t2["y_vec"], tree["y_vec"]) and np.array_equal(t2["xz_end_vec"], tree["xz_end_vec"])
    back.close()
    other.close()
    apr.close()
