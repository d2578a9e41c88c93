"""Pins the oracle's image helpers: pre-blur taps and filter
(src/GaussianBlurCUDA.cu:39-130), window sums (src/array2d_refImg.cu:11-131)
and the transpose (src/cudaTranspose.h:8-71)."""
import numpy as np
import pytest

from oracle import pmx_oracle as O

GAUSS, BOX = 1, 2


@pytest.mark.parametrize("sigma,width", [(0.5, 5), (1.0, 5), (1.3, 7), (2.0, 11), (3.0, 15),
                                         (8.9, 45), (10.0, 45)])
def test_blur_tap_widths_and_gaussian_closed_form(sigma, width):
    t = O.blur_taps(sigma, GAUSS)
    assert len(t) == width
    i = np.arange(width) - width // 2
    g = np.exp(-(i * i) / (2.0 * sigma * sigma))
    np.testing.assert_allclose(t, g / g.sum(), rtol=1e-15, atol=0)
    assert abs(t.sum() - 1.0) < 1e-15


@pytest.mark.parametrize("sigma", [1.0, 2.0, 20.0])
def test_box_taps_are_uniform(sigma):
    t = O.blur_taps(sigma, BOX)
    assert np.all(t == 1.0 / len(t))


def test_impulse_response_recovers_taps():
    for mode, sigma in [(GAUSS, 1.0), (GAUSS, 2.0), (BOX, 1.0)]:
        t = O.blur_taps(sigma, mode)
        img = np.zeros((41, 43))
        img[20, 21] = 1.0
        out = O.blur(img, sigma, mode)
        r = len(t) // 2
        np.testing.assert_allclose(out[20 - r:21 + r, 21 - r:22 + r], np.outer(t, t), atol=1e-16)
        assert abs(out.sum() - 1.0) < 1e-13


def test_blur_constant_fixed_point_and_clamp():
    img = np.full((17, 23), 0.37)
    np.testing.assert_allclose(O.blur(img, 1.7, GAUSS), img, atol=1e-15)
    # clamp-to-edge: a ramp in x stays a ramp in the interior and bends at the border
    ramp = np.tile(np.arange(23, dtype=np.float64), (17, 1))
    out = O.blur(ramp, 1.0, GAUSS)
    np.testing.assert_allclose(out[:, 3:-3], ramp[:, 3:-3], atol=1e-12)
    t = O.blur_taps(1.0, GAUSS)
    assert abs(out[5, 0] - (t[3] * 1 + t[4] * 2)) < 1e-12   # taps left of x=0 read x=0 (value 0)


def _brute_box(img, hw):
    H, W = img.shape
    a = np.zeros_like(img)
    b = np.zeros_like(img)
    for y in range(H):
        for x in range(W):
            win = img[max(0, y - hw):y + hw + 1, max(0, x - hw):x + hw + 1]
            a[y, x] = win.sum()
            b[y, x] = (win * win).sum()
    return a, b


@pytest.mark.parametrize("hw", [1, 2, 5, 7])
def test_box_sums_brute_force(hw):
    rs = np.random.default_rng(hw)
    img = rs.random((13, 19))
    a, b = O.box_sums(img, hw)
    ea, eb = _brute_box(img, hw)
    np.testing.assert_allclose(a, ea, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(b, eb, rtol=1e-13, atol=1e-13)


def test_box_sums_constant_counts_in_bounds():
    hw, c = 3, 0.25
    a, b = O.box_sums(np.full((11, 9), c), hw)
    ny = np.array([min(y + hw, 10) - max(y - hw, 0) + 1 for y in range(11)])
    nx = np.array([min(x + hw, 8) - max(x - hw, 0) + 1 for x in range(9)])
    np.testing.assert_allclose(a, c * np.outer(ny, nx), rtol=1e-15)
    np.testing.assert_allclose(b, c * c * np.outer(ny, nx), rtol=1e-15)


@pytest.mark.parametrize("shape", [(1, 1), (31, 33), (33, 31), (1, 17), (16, 1)])
def test_transpose_is_permutation(shape):
    img = np.arange(shape[0] * shape[1], dtype=np.float64).reshape(shape)
    assert np.array_equal(O.transpose(img), img.T)


def test_no_blur_mode_is_the_identity():
    """Blur mode 0 leaves the image as it is (the reference never calls its
    filter, src/patchMatch.cu:150-205), for any sigma."""
    img = np.random.default_rng(3).uniform(0, 1, (23, 31))
    for sigma in (0.5, 1.0, 4.0):
        assert np.array_equal(O.blur(img, sigma, 0), img)
