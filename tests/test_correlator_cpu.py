"""Pins of the FFT imaging-condition restatement (oracle/correlator.py, bgprtm/ImagingCorrelator.cpp) by
closed forms: plane waves through the unit derivatives, the inversion and kfilter imaging conditions
and costerm; the identity sum_a H_a^2 = -I away from k = 0; skew-adjointness of H_a."""
import numpy as np
import pytest

from oracle import correlator as oc
from tests.helpers import bandlimited_field

D = (10.0, 12.5, 7.5)   # dx, dy, dz


def _grid(shape):
    ny, nx, nz = shape
    y, x, z = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    return y, x, z


def _plane(shape, m):
    """cos(theta) and sin(theta) of the grid-periodic wave with integer wavenumbers m = (my, mx, mz), and its
    physical wavevector (ky, kx, kz) built from the FP32 dk of assignW."""
    ny, nx, nz = shape
    y, x, z = _grid(shape)
    th = 2 * np.pi * (m[0] * y / ny + m[1] * x / nx + m[2] * z / nz)
    dk = [float(np.float32(2.0 * np.pi / float(np.float32(dd) * np.float32(n))))
          for dd, n in ((D[1], ny), (D[0], nx), (D[2], nz))]
    k = np.array([m[0] * dk[0], m[1] * dk[1], m[2] * dk[2]])
    return np.cos(th), np.sin(th), k


def _smooth(shape, seed, cut=0.3):
    return bandlimited_field(shape, seed, cut)


def test_assign_w_convention():
    w = oc.assign_w(8, 2.0)
    dk = np.float32(2 * np.pi / 16.0)
    assert np.array_equal(w, np.array([0, -1, -2, -3, -4, 3, 2, 1], np.float32) * dk)


@pytest.mark.parametrize("m", [(1, 2, 3), (0, 3, 1), (2, 0, 0), (1, 1, 5)])
def test_unit_derivative_of_plane_wave(m):
    shape = (8, 10, 12)
    c, s, k = _plane(shape, m)
    kk = np.linalg.norm(k)
    # H_a cos(theta) = (k_a/|k|) sin(theta) for the reference's sign of the wavenumbers
    for axis, ka in ((0, k[2]), (1, k[1]), (2, k[0])):
        got = oc.unit_derivative(c, axis, D)
        np.testing.assert_allclose(got, (ka / kk) * s, atol=1e-12)


@pytest.mark.parametrize("m", [(1, 2, 3), (0, 3, 1), (2, 1, 4)])
def test_inversion_of_plane_wave_is_cos_2theta(m):
    shape = (8, 10, 12)
    c, s, k = _plane(shape, m)
    img, fwi, _ = oc.run(1, c, c, D, image=np.zeros(shape), img_fwi=np.zeros(shape))
    np.testing.assert_allclose(img, c * c - s * s, atol=1e-12)     # cos^2 - sum (k_a/|k|)^2 sin^2
    np.testing.assert_allclose(fwi, np.ones(shape), atol=1e-12)    # cos^2 + sin^2
    # ALLIN3: the raw image |S R| first, then the same inversion terms
    img3, fwi3, raw = oc.run(3, c, c, D, image=np.zeros(shape), img_fwi=np.zeros(shape), img_raw=np.zeros(shape))
    np.testing.assert_allclose(raw, c * c, atol=1e-12)
    np.testing.assert_allclose(img3, img, atol=1e-12)


def test_inversion_2d_has_no_y_term():
    shape = (1, 16, 12)
    c, s, k = _plane(shape, (0, 3, 2))
    img, _, _ = oc.run(1, c, c, D, image=np.zeros(shape))
    np.testing.assert_allclose(img, c * c - s * s, atol=1e-12)


def test_sum_of_squared_unit_derivatives_is_minus_identity():
    shape = (12, 14, 16)
    x = _smooth(shape, 3)
    acc = sum(oc.unit_derivative(oc.unit_derivative(x, a, D), a, D) for a in (0, 1, 2))
    np.testing.assert_allclose(acc, -x, atol=1e-10 * np.abs(x).max())


def test_unit_derivative_is_skew_adjoint():
    shape = (10, 12, 14)
    u, v = _smooth(shape, 5), _smooth(shape, 6)
    for a in (0, 1, 2):
        lhs = np.vdot(oc.unit_derivative(u, a, D), v)
        rhs = -np.vdot(u, oc.unit_derivative(v, a, D))
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs) + 1e-12


def test_kfilter_and_costerm_of_plane_wave():
    shape = (8, 10, 12)
    c, s, k = _plane(shape, (1, 2, 3))
    kk = np.linalg.norm(k)
    np.testing.assert_allclose(oc.kfilter(c, D, 0.1 * kk), c / kk, atol=1e-12)
    np.testing.assert_allclose(oc.kfilter(c, D, 2.0 * kk), c / (2.0 * kk), atol=1e-12)   # clipped at kmin
    np.testing.assert_allclose(oc.kfilter(np.full(shape, 3.0), D, 0.5), np.full(shape, 6.0), atol=1e-12)
    np.testing.assert_allclose(oc.costerm(c, D), kk * kk * c, atol=1e-9 * kk * kk)
    # inversion2: imgFWI takes S R, the image K S K R
    img, fwi, _ = oc.run(2, c, c, D, kmin=0.1 * kk, image=np.zeros(shape), img_fwi=np.zeros(shape))
    np.testing.assert_allclose(fwi, c * c, atol=1e-12)
    np.testing.assert_allclose(img, c * c / (kk * kk), atol=1e-12)
    np.testing.assert_allclose(oc.sinterm(fwi, img), 2 * c * c - 0.5 * c * c / (kk * kk), atol=1e-12)
