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


# ---- doubled imaging grids ------------------------------------------------------------------------------

def _gauss(shape, c, sigma, shift=(0.0, 0.0, 0.0)):
    """exp(-|p + shift - c|^2 / (2 sigma^2)) sampled at the grid points p (y, x, z)."""
    y, x, z = _grid(shape)
    r2 = (y + shift[0] - c[0]) ** 2 + (x + shift[1] - c[1]) ** 2 + (z + shift[2] - c[2]) ** 2
    return np.exp(-r2 / (2.0 * sigma * sigma))


def test_fft_size_is_smallest_5_smooth():
    smooth = [m for m in range(1, 400) if all(p in (2, 3, 5) for p in _factors(m))]
    for n in (1, 7, 17, 97, 121, 250, 251):
        assert oc.fft_size(n) == min(m for m in smooth if m >= n)


def _factors(m):
    out, p = [], 2
    while m > 1:
        while m % p == 0:
            out.append(p)
            m //= p
        p += 1
    return out


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_half_shift_of_gaussian_is_analytic(axis):
    """SHIFTHALF along each axis of a band-limited Gaussian equals the Gaussian at p + 1/2 on that axis (the
    position interleave gives the copy); a wrong sign would give p - 1/2, a lost factor 2 a whole sample."""
    shape = (24, 28, 32)
    c = (11.3, 13.7, 15.2)
    g = _gauss(shape, c, 2.5)
    sh = [0.0, 0.0, 0.0]
    sh[{0: 2, 1: 1, 2: 0}[axis]] = 0.5
    got = oc.line_filter(g, axis, 0)
    assert np.abs(got - _gauss(shape, c, 2.5, sh)).max() < 2e-6
    sh[{0: 2, 1: 1, 2: 0}[axis]] = -0.5
    assert np.abs(got - _gauss(shape, c, 2.5, sh)).max() > 1e-2


def test_high_tape_passes_low_and_removes_near_nyquist():
    shape = (1, 6, 128)
    g = _gauss(shape, (0, 2.5, 63.3), 8.0)
    assert np.abs(oc.line_filter(g, 0, 1, 0.7) - g).max() < 1e-6     # spectrum well inside 0.7 Nyquist
    alt = g * (-1.0) ** np.arange(128)                                 # the same envelope at Nyquist
    out = oc.line_filter(alt, 0, 1, 0.7)
    assert np.linalg.norm(out) < 0.1 * np.linalg.norm(alt)
    assert np.abs(oc.line_filter(alt, 0, 1, 1.0) - alt).max() < 1e-9  # h = 1 is the identity


def test_double_copy_order_matches_reference_indices():
    """ivz=1, ivx=2, ivy=3, ivzx=4, ivzy=5, ivxy=6, ivzxy=7 in 3D; ivz=1, ivx=2, ivzx=3 in 2D; ivz=1 alone
    (bgprtm/ImagingCorrelator.cpp:58-75); zx from z, zy from z, xy from x, zxy from zx (:515-524)."""
    c = oc.double_copies(2, 2, 2)
    assert [b for b, _, _ in c] == [0, 1, 2, 4, 3, 5, 6, 7]
    assert [(s, a) for _, s, a in c[1:]] == [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2), (4, 2)]
    assert [b for b, _, _ in oc.double_copies(2, 2, 1)] == [0, 1, 2, 3]
    assert [b for b, _, _ in oc.double_copies(2, 1, 1)] == [0, 1]


@pytest.mark.parametrize("dbl,shape", [((2, 2, 2), (5, 6, 7)), ((2, 2, 1), (1, 6, 7)), ((2, 1, 1), (1, 6, 7))])
def test_interleave_places_copies_at_half_positions(dbl, shape):
    """Sub-image i holding f at the copy's shifted position interleaves to f sampled on the doubled grid."""
    f = lambda y, x, z: 1000.0 * y + 10.0 * x + 0.1 * z + 0.01   # injective on the half grid
    y, x, z = _grid(shape)
    sub = np.stack([f(y + ((b >> 2) & 1) / 2, x + ((b >> 1) & 1) / 2, z + (b & 1) / 2)
                    for b, _, _ in oc.double_copies(*dbl)])
    out = oc.interleave(sub, dbl)
    dz, dx, dy = dbl
    Y, X, Z = _grid(out.shape)
    assert np.array_equal(out, f(Y / dy, X / dx, Z / dz))


def test_interp_then_interleave_upsamples_band_limited_field():
    """interleave(interp(g)) of a smooth Gaussian is the Gaussian sampled on the half-spacing grid."""
    shape = (28, 30, 32)
    c = (13.4, 14.9, 15.6)
    g = _gauss(shape, c, 2.5)
    out = oc.interleave(np.stack(oc.interp(g, (2, 2, 2))), (2, 2, 2))
    Y, X, Z = _grid(out.shape)
    want = np.exp(-((Y / 2 - c[0]) ** 2 + (X / 2 - c[1]) ** 2 + (Z / 2 - c[2]) ** 2) / (2 * 2.5 ** 2))
    assert np.abs(out - want).max() < 1e-5


def test_interp_illum_averages_linear_fields_exactly():
    """Averaging neighbours reproduces a linear field at the half positions away from the clamped last plane;
    constants stay constant on every layout; in 2D with z only doubled the reference's later writes at the
    shared x index stand, so the even z samples carry the x-neighbour average."""
    shape = (6, 7, 8)
    y, x, z = _grid(shape)
    lin = (3.0 * y + 2.0 * x + 0.5 * z).astype(np.float32)
    out = oc.interp_illum(lin, (2, 2, 2))
    Y, X, Z = _grid(out.shape)
    inner = (Y < 2 * shape[0] - 1) & (X < 2 * shape[1] - 1) & (Z < 2 * shape[2] - 1)
    assert np.array_equal(out[inner], (3.0 * Y / 2 + 2.0 * X / 2 + 0.5 * Z / 2)[inner].astype(np.float32))
    for dbl, sh in (((2, 2, 2), (4, 5, 6)), ((2, 2, 1), (1, 5, 6)), ((2, 1, 1), (1, 5, 6))):
        assert np.array_equal(oc.interp_illum(np.full(sh, 1.75, np.float32), dbl), np.full(
            (sh[0] * dbl[2], sh[1] * dbl[1], sh[2] * dbl[0]), 1.75, np.float32))
    x2 = np.broadcast_to(np.arange(5, dtype=np.float32)[None, :, None], (1, 5, 6)).copy()
    out = oc.interp_illum(x2, (2, 1, 1))
    assert np.array_equal(out[0, :4, 0::2], np.broadcast_to((np.arange(4) + 0.5)[:, None], (4, 6)).astype(np.float32))


@pytest.mark.parametrize("option", [0, 1])
def test_doubled_run_reduces_to_plain_run(option):
    shape = (6, 8, 10)
    s, r = bandlimited_field(shape, 5), bandlimited_field(shape, 6)
    _, img, fwi = oc.run_double(option, s, r, D, (1, 1, 1), (0.9, 0.7), np.zeros((1,) + shape), np.zeros((1,) + shape))
    wi, wf, _ = oc.run(option, s, r, D, 0.0, np.zeros(shape), np.zeros(shape))
    assert np.allclose(img[0], wi, atol=1e-12)
    if option == 1:
        assert np.allclose(fwi[0], wf, atol=1e-12)


def test_doubled_image1_is_product_of_upsampled_fields():
    """IMAGE1 on the doubled grid with no high cut: interleaved sub-images = g1 g2 at the half positions."""
    shape = (28, 30, 32)
    c1, c2 = (13.3, 14.2, 15.7), (14.1, 15.4, 16.2)
    g1, g2 = _gauss(shape, c1, 2.2), _gauss(shape, c2, 2.6)
    nl = len(oc.double_copies(2, 2, 2))
    _, img, _ = oc.run_double(0, g1, g2, D, (2, 2, 2), (1.0, 1.0), np.zeros((nl,) + shape), None)
    out = oc.interleave(img, (2, 2, 2))
    Y, X, Z = _grid(out.shape)
    h = lambda c, s: np.exp(-((Y / 2 - c[0]) ** 2 + (X / 2 - c[1]) ** 2 + (Z / 2 - c[2]) ** 2) / (2 * s * s))
    assert np.abs(out - h(c1, 2.2) * h(c2, 2.6)).max() < 1e-5
