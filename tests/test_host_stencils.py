"""Host-only parts of libaprh (no GPU needed): stencil preparation.

* get_downsampled_stencils (APRStencil.hpp:12-97) bit-equal to the reference build and
  to the oracle, normalised stacks sum to one;
* create_gaussian_filter (APRStencil.hpp:173-228) bit-equal to the reference build and
  within float rounding of the double-precision closed form;
* create_mean_filter (APRStencil.hpp:236-244) is 1/n.
"""
import numpy as np
import pytest

import libapr_amd as A
import oracle_lib as O


def lib_or_skip():
    try:
        A.lib()
    except RuntimeError as e:
        pytest.skip(str(e))


STENCILS = {
    "mean3": lambda: np.full((3, 3, 3), 1.0 / 27, np.float32),
    "gauss5": lambda: O.gaussian_stencil((1.0, 1.0, 1.0), (5, 5, 5), True),
    "aniso": lambda: O.gaussian_stencil((1.0, 1.5, 2.0), (3, 5, 7), False),
    "gauss9": lambda: O.gaussian_stencil((1.0, 1.0, 2.0), (9, 9, 9), True),
    "flat2d": lambda: O.gaussian_stencil((1.0, 2.0, 1.0), (5, 7, 1), True),
}


@pytest.mark.parametrize("kind", sorted(STENCILS))
@pytest.mark.parametrize("normalize", [False, True])
def test_downsampled_stencils_equal_oracle_and_reference(kind, normalize):
    lib_or_skip()
    st = STENCILS[kind]()
    got = A.APRStencil.get_downsampled_stencils(st, 7, normalize)
    exp = O.downsampled_stencils(st, 7, normalize, "oracle")
    assert len(got) == len(exp)
    for a, b in zip(got, exp):
        assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))
        if normalize:
            assert abs(float(a.astype(np.float64).sum()) - 1.0) < 1e-5
    if O.load("ref") is not None:
        for a, b in zip(got, O.downsampled_stencils(st, 7, normalize, "ref")):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_downsampled_stencil_mean_closed_form():
    """a mean stencil downsampled by 2^d stays a (clipped) mean: every tap of the 3^3
    result is the overlap fraction of the input box"""
    lib_or_skip()
    st = np.full((3, 3, 3), 1.0 / 27, np.float32)
    d1 = A.APRStencil.get_downsampled_stencils(st, 2, False)[1]
    # each input tap spans 2 units starting at 1 + i (offset (3 // 2) * 2 - 3 // 2 = 1); output
    # cells span [2d, 2d + 2): per-axis overlap sums 1, 4, 1; then / 2^3
    w = np.array([1.0, 4.0, 1.0])
    exp = np.einsum("i,j,k->ijk", w, w, w) / 27.0 / 8.0
    np.testing.assert_allclose(d1, exp, rtol=1e-6)


@pytest.mark.parametrize("sigma,size", [((1.0, 1.0, 1.0), (5, 5, 5)), ((1.0, 1.5, 2.0), (3, 5, 7)),
                                        ((0.7, 2.5, 1.3), (9, 1, 5))])
@pytest.mark.parametrize("normalize", [False, True])
def test_gaussian_filter(sigma, size, normalize):
    lib_or_skip()
    g = A.APRStencil.create_gaussian_filter(sigma, size, normalize)
    assert g.shape == (size[2], size[1], size[0])
    if O.load("ref") is not None:
        r = O.gaussian_stencil_lib(sigma, size, normalize, "ref")
        assert np.array_equal(g.view(np.uint32), r.view(np.uint32))
    # closed form in double
    def g1(s, n):
        if n <= 1:
            return np.ones(1)
        i = np.arange(-(n // 2), n // 2 + 1, dtype=np.float64)
        return np.exp(-i * i / (2 * s * s)) / (s * np.sqrt(2 * np.pi))
    e = np.einsum("k,j,i->kji", g1(sigma[2], size[2]), g1(sigma[1], size[1]), g1(sigma[0], size[0]))
    if normalize:
        e = e / e.sum()
    np.testing.assert_allclose(g, e, rtol=2e-6)


def test_mean_filter():
    lib_or_skip()
    m = A.APRStencil.create_mean_filter((3, 5, 7))
    assert m.shape == (7, 5, 3)
    assert np.all(m == np.float32(1.0 / 105.0))


def test_stencil_argument_errors():
    lib_or_skip()
    with pytest.raises(A.AprhError):
        A.APRStencil.get_downsampled_stencils(np.ones((2, 3, 3), np.float32), 3, False)   # even size
    with pytest.raises(A.AprhError):
        A.APRStencil.get_downsampled_stencils(np.ones((3, 3, 3), np.float32), 0, False)
