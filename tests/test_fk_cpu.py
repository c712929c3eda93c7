"""Pins of the FK pre-processing restatement (oracle/fk.py, bgprtm/Receiver.cpp:851-1010) by closed forms on
plane waves that are periodic over the transform box: the obliquity factor of true amplitude, the deghost
operator's 90-degree turn with 0.5/sin(kz z), its zero-frequency threshold and low cut, the removal of
evanescent waves and of the zero frequency, and the unnormalised round trip (gain sqrt(nx ny ntfft))."""
import numpy as np
import pytest

from oracle import fk
from oracle import pipeline

DX, DY, DT, V = 12.5, 10.0, 0.004, 1500.0


def _plane(ny, nx, nt, my, mx, mt):
    y, x, t = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nt), indexing="ij")
    return 2 * np.pi * (my * y / ny + mx * x / nx - mt * t / nt)


def _k(ny, nx, nt, my, mx, mt):
    f = np.float32
    dkx = float(f(2 * np.pi / float(f(DX) * f(nx))))
    dky = float(f(2 * np.pi / float(f(DY) * f(ny))))
    dwv = float(f(2 * np.pi / float(f(f(DT) * f(nt)) * f(V))))
    return mx * dkx, my * dky, mt * dwv


@pytest.mark.parametrize("m", [(1, 2, 9), (0, 3, 12), (2, 1, 20)])
@pytest.mark.parametrize("trueamp", [0, 1])
def test_true_amplitude_plane_wave(m, trueamp):
    ny, nx, nt = 8, 10, 48
    phi = _plane(ny, nx, nt, *m)
    kx, ky, k = _k(ny, nx, nt, *m)
    cos = np.sqrt(1 - (kx * kx + ky * ky) / (k * k))
    assert cos > 0.1
    out = fk.fk_filter(np.cos(phi), DX, DY, DT, V, trueamp, 0.0, 0, -1)
    gain = np.sqrt(ny * nx * nt) * (cos if trueamp else 1.0)
    np.testing.assert_allclose(out, gain * np.cos(phi), atol=1e-9 * gain)


@pytest.mark.parametrize("sign", [-1, 1])
def test_deghost_plane_wave(sign):
    ny, nx, nt = 8, 10, 48
    m = (1, 2, 9)
    phi = _plane(ny, nx, nt, *m)
    kx, ky, k = _k(ny, nx, nt, *m)
    cos = np.sqrt(1 - (kx * kx + ky * ky) / (k * k))
    for z, thr in ((8.0, 0.1), (30.0, 0.1), (8.0, 0.9)):
        kzdz = k * cos * z
        g = 1.0 if kzdz >= np.pi / 2 else (1 / thr if np.sin(kzdz) <= thr else 1 / np.sin(kzdz))
        out = fk.fk_filter(np.cos(phi), DX, DY, DT, V, 1, z, 1, sign, thr=thr)
        want = sign * np.sqrt(ny * nx * nt) * cos * 0.5 * g * np.sin(phi)     # i*sign turns cos into sign*sin
        np.testing.assert_allclose(out, want, atol=1e-9 * np.abs(want).max())


def test_low_cut_and_zero_frequency():
    ny, nx, nt = 6, 8, 64
    # frequencies mt = 3 (below the cut) and mt = 12 (inside the ramp); a constant offset (w = 0)
    f = lambda mt: np.cos(_plane(ny, nx, nt, 0, 0, mt))                     # vertical waves: cos = 1
    flcut, flpass = 4.0 / (DT * nt), 16.0 / (DT * nt)                        # iw_lc = 4, iw_lp = 16
    out = fk.fk_filter(f(3) + f(12) + 0.7, DX, DY, DT, V, 0, 5.0, 1, 1, flcut=flcut, flpass=flpass)
    k = 12 * float(np.float32(2 * np.pi / float(np.float32(np.float32(DT) * np.float32(nt)) * np.float32(V))))
    g = 1.0 / np.sin(k * 5.0)
    taper = 0.5 * (1 - np.cos(np.pi * (12 - 4) / (16 - 4 + 1)))
    want = np.sqrt(ny * nx * nt) * 0.5 * g * taper * np.sin(_plane(ny, nx, nt, 0, 0, 12))
    np.testing.assert_allclose(out, want, atol=1e-9 * np.abs(want).max())


def test_evanescent_waves_are_removed():
    ny, nx, nt = 8, 10, 48
    m = (3, 4, 2)                   # horizontal wavenumber far above w / V
    kx, ky, k = _k(ny, nx, nt, *m)
    assert kx * kx + ky * ky > k * k
    out = fk.fk_filter(np.cos(_plane(ny, nx, nt, *m)), DX, DY, DT, V, 1, 0.0, 0, -1)
    assert np.abs(out).max() < 1e-9


def test_xy_spread_of_one_trace_is_separable():
    nx, ny, nz = 30, 26, 20
    tab = pipeline.spread_tables(nx, ny, nz, np.array([[14.3, 11.6, 8.0]], np.float32), 5, 0.3, 0.3)
    tr = np.random.default_rng(1).standard_normal((1, 16))
    d = fk.xy_spread(tr, tab, nx, ny, 24)
    kx = tab["kx"][0, :tab["lx"][0]].astype(np.float64)
    ky = tab["ky"][0, :tab["ly"][0]].astype(np.float64)
    blk = d[tab["offy"][0]:tab["offy"][0] + ky.size, tab["offx"][0]:tab["offx"][0] + kx.size, :16]
    np.testing.assert_allclose(blk, ky[:, None, None] * kx[None, :, None] * tr[0][None, None, :], rtol=1e-12)
    assert np.abs(d).sum() == pytest.approx(np.abs(blk).sum())
    assert np.all(d[:, :, 16:] == 0)
