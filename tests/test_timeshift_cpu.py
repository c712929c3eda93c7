"""Oracle pinning of the FFT time shift (FreqUtil::timeshift_in/out, bgprtm/FreqUtil.cpp:24-39) by what
the transform fixes: an integer shift is the circular shift of the padded trace, shifts compose, a
zero shift is the identity, and a fractional shift of a periodic band-limited signal is the signal
evaluated at the shifted times.  The padded length (smallest 2^a 3^b 5^c >= nt) stands in for the
unvendored libCommon::padfft: parity unpinned for that choice, pinned here only by its definition."""
import numpy as np
import pytest

import oracle


def test_fft_size():
    assert [oracle.fft_size(n) for n in (1, 2, 7, 11, 13, 601, 2999, 3000, 3001)] == \
        [1, 2, 8, 12, 15, 625, 3000, 3000, 3072]


@pytest.mark.parametrize("k", [0, 1, 5, 37, -4])
def test_integer_shift_is_circular(k):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(250)
    n = oracle.fft_size(250)
    pad = np.zeros(n)
    pad[:250] = x
    assert np.allclose(oracle.timeshift(x, k, n), np.roll(pad, k)[:250], atol=1e-12)


def test_shifts_compose_and_zero_is_identity():
    # a band-limited pulse: the Nyquist bin (whose imaginary part the real inverse drops) is empty
    t = np.arange(300)
    x = np.exp(-((t - 120) / 12.0) ** 2) * np.cos(0.3 * (t - 120))
    n = oracle.fft_size(300)
    assert np.allclose(oracle.timeshift(x, 0.0, n), x, atol=1e-12)
    # compose on the padded length, so no samples are lost between the two shifts
    pad = np.zeros(n)
    pad[:300] = x
    a = oracle.timeshift(oracle.timeshift(pad, 2.3, n), 4.45, n)
    assert np.allclose(a, oracle.timeshift(pad, 6.75, n), atol=1e-10)


def test_fractional_shift_of_a_periodic_sinusoid():
    n = 240
    t = np.arange(n)
    f = 7 / n  # seven periods over the padded length: exactly representable on the grid
    x = np.cos(2 * np.pi * f * t + 0.3)
    got = oracle.timeshift(x, 3.37, n)
    assert np.allclose(got, np.cos(2 * np.pi * f * (t - 3.37) + 0.3), atol=1e-10)
