"""APR files (libaprh_io.so, APRFile): CPU checks.

* The reference's own test files (tests/golden/<case>.apr, the sphere.apr files of its
  test/files/Apr/<case>/): their linear access is the structure the oracle computes from the
  reference test's saved gradient and scale (get_apr_custom_grad_scale, as
  APRTest.cpp:2816-2825 compares), and their particle intensities are the oracle's sampling
  of the original image (APRTest.cpp:2827-2836).
* Round trips: structure and u16 / f32 particles written and read back unchanged, several
  time points and channels in one file, the tree group, overwrite rules; the written layout
  (dataset names, types, blosc filter 32001 and its parameters) is the reference file's, read
  with HDF5's own h5ls where it is installed."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import libapr_amd as A
import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")

pytestmark = pytest.mark.skipif(not os.path.exists(A.IO_LIB_PATH), reason="libaprh_io.so not built")


def _fixture(case):
    d = np.load(os.path.join(GOLD, case + ".npz"))
    return {k: d[k] for k in d.files}


@pytest.mark.parametrize("case", ["sphere_120", "sphere_diff_dims"])
def test_reference_file_matches_oracle(case):
    f = _fixture(case)
    af = A.APRFile()
    af.open(os.path.join(GOLD, case + ".apr"), "READ")
    dims, xze, yv = af.read_structure()
    parts = af.read_particles("particle_intensities")
    af.close()
    img = np.ascontiguousarray(f["original"])
    Z, X, Y = img.shape
    assert dims == (Y, X, Z)
    assert yv.size == int(f["attr_total_number_particles"])
    par = O.params(Ip_th=0.0, rel_error=float(f["attr_rel_error"]), lambda_=float(f["attr_lambda"]),
                   sigma_th_max=max(float(f["attr_sigma_th_max"]), 0.1), sigma_th=max(float(f["attr_sigma_th"]), 0.1),
                   grad_th=float(f["attr_grad_th"]))
    inten = np.ones(f["scale_saved"].shape, np.float32)
    o = O.apr_from_grad_lis(f["gradient_saved"], f["scale_saved"], inten, (Y, X, Z), par, 0.0)
    assert np.array_equal(o.xz_end_vec, xze)
    assert np.array_equal(o.y_vec, yv)
    assert np.array_equal(o.sample(img), parts)


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    af = A.APRFile()
    src = os.path.join(GOLD, "sphere_diff_dims.apr")
    af.open(src, "READ")
    dims, xze, yv = af.read_structure()
    parts = af.read_particles("particle_intensities")
    af.close()
    pf = rng.standard_normal(yv.size).astype(np.float32)
    path = str(tmp_path / "rt.apr")
    af.open(path, "WRITE")
    par = A.APRParameters(rel_error=0.2, lambda_=2.5, sigma_th=30, grad_th=7)
    af.write_structure(dims, xze, yv, parameters=par)
    af.write_particles("particles", parts)
    af.write_particles("floats", pf)
    af.write_structure(dims, xze[:1] * 0, yv[:0], t=3, channel_name="ch")   # an empty APR at another time point
    af.close()
    af.open(path, "READWRITE")
    af.write_particles("particles", parts[::-1].copy())            # same size and type: overwritten
    with pytest.raises(A.AprhError):
        af.write_particles("particles", parts[:10].copy())          # another size: refused
    af.close()
    af.open(path, "READ")
    d2, x2, y2 = af.read_structure()
    assert d2 == dims and np.array_equal(x2, xze) and np.array_equal(y2, yv)
    assert np.array_equal(af.read_particles("particles"), parts[::-1])
    got = af.read_particles("floats")
    assert got.dtype == np.float32 and np.array_equal(got.view(np.uint32), pf.view(np.uint32))
    assert af.read_info(t=3, channel_name="ch")[1] == 0
    with pytest.raises(A.AprhError):
        af.read_structure(t=1)
    with pytest.raises(A.AprhError):
        af.read_particles("missing")
    af.close()
    with pytest.raises(A.AprhError):
        af.open(str(tmp_path / "none.apr"), "READ")


def test_written_layout_is_the_reference_layout(tmp_path):
    h5ls = shutil.which("h5ls") or ("/opt/conda/bin/h5ls" if os.path.exists("/opt/conda/bin/h5ls") else None)
    if h5ls is None:
        pytest.skip("h5ls not installed")
    af = A.APRFile()
    af.open(os.path.join(GOLD, "sphere_120.apr"), "READ")
    dims, xze, yv = af.read_structure()
    parts = af.read_particles("particle_intensities")
    af.close()
    path = str(tmp_path / "w.apr")
    af.open(path, "WRITE")
    af.write_structure(dims, xze, yv)
    af.write_particles("particle_intensities", parts)
    af.close()

    def layout(p):
        out = subprocess.run([h5ls, "-r", "-v", p], capture_output=True, text=True).stdout
        items, cur = {}, None
        for line in out.splitlines():
            if "Dataset" in line:
                cur = line.split()[0]
                items[cur] = []
            elif cur and ("Filter-0" in line or line.strip().startswith("Type:")):
                items[cur].append(line.split(":", 1)[1].strip())
        return items

    ref, ours = layout(os.path.join(GOLD, "sphere_120.apr")), layout(path)
    assert set(ref) == set(ours)
    for k in ref:
        assert ref[k] == ours[k], (k, ref[k], ours[k])
