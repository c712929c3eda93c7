"""FFT time shift on the GPU (hipFFT; FreqUtil::timeshift_in/out, bgprtm/FreqUtil.cpp:24-39) against the
FP64 oracle, and a shot with a delayed source array (Propagator::setupReceiver,
bgprtm/Propagator.cpp:783-811) against the oracle shot."""
import numpy as np
import pytest

import oracle
from oracle import pipeline
from konglong_amd._lib import DeviceArray, check, lib, fptr, ENGINE_FD, ENGINE_CFD, ISO, VTI, SNAP_FP32
from konglong_amd.propagator import Propagator, shot_config
from tests.helpers import rel_l2, smooth_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nt", [600, 2999, 1001])
def test_timeshift_matches_oracle(klrtm, nt):
    print(f"[timeshift] nt={nt}", flush=True)
    n = int(lib().klrtm_fft_size(nt))
    assert n == oracle.fft_size(nt)
    w = pipeline.ricker(nt, 1e-3, 15.0)
    shifts = np.array([0.0, 1.0, 2.5, -3.25, 17.7, 250.0], np.float32)
    din, dout = DeviceArray.from_host(w), DeviceArray((shifts.size, nt))
    check(lib().klrtm_timeshift(din.ptr, nt, n, shifts.size, fptr(shifts), dout.ptr))
    got = dout.download()
    for s, row in zip(shifts, got):
        want = oracle.timeshift(w, float(s), n)
        assert rel_l2(row, want) < 2e-6, s
    assert np.allclose(got[0], w, atol=1e-6 * np.abs(w).max())


def test_timeshift_rejects_bad_sizes(klrtm):
    d = DeviceArray((64,))
    sh = np.zeros(1, np.float32)
    assert lib().klrtm_timeshift(d.ptr, 64, 32, 1, fptr(sh), d.ptr) != 0   # ntfft < nt
    assert lib().klrtm_timeshift(d.ptr, 64, 64, 0, fptr(sh), d.ptr) != 0   # no traces


@pytest.mark.parametrize("case", [(ENGINE_FD, 8, ISO), (ENGINE_CFD, 10, VTI)], ids=["FD8-ISO", "CFD10-VTI"])
def test_source_array_shot_matches_oracle(klrtm, case):
    engine, order, model = case
    print(f"[source-array] {case}", flush=True)
    ny, nx, nz = 32, 36, 40
    d, dt, nt, nb = 10.0, 1e-3, 130, 8
    vel = smooth_field((ny, nx, nz), 1800.0, 2600.0, 91)
    eps = smooth_field((ny, nx, nz), 0.0, 0.2, 92) if model == VTI else None
    dl = smooth_field((ny, nx, nz), 0.0, 0.1, 93) if model == VTI else None
    rcv = np.array([[x, y, 10.0] for y in range(2, ny - 2, 3) for x in range(2, nx - 2, 3)], np.float32)
    srcs = np.array([[12.0, 14.0, 13.0], [18.5, 15.0, 13.0], [24.0, 16.2, 14.0]], np.float32)
    delays = np.array([0.0, 0.0043, 0.0115], np.float32)
    k1, k2 = (0.3, 0.3) if engine == ENGINE_FD else (0.40, 0.48)
    cfg = shot_config(nx, ny, nz, d, dt, nt, engine, order, model, nbnd=nb, vmax=2600.0, src_xyz=tuple(srcs[0]),
                      f_peak=15.0, spread=6, khpass=k1, khcut=k2, nrcv=rcv.shape[0], snap_every=2,
                      snap_format=SNAP_FP32, tmute_frac=0.1)
    p = Propagator(cfg)
    p.set_model(vel, eps, dl)
    p.set_receivers(rcv)
    p.set_source_array(srcs, delays)
    o = pipeline.Shot(nx, ny, nz, d, dt, nt, engine, order, model, [nb] * 6, 2600.0, tuple(srcs[0]), 15.0, 6, k1, k2,
                      rcv, snap_every=2, tmute_frac=0.1, src_array=(srcs, delays))
    o.set_model(vel, eps, dl)
    p.set_source_wavelet(o.wavelet)   # re-shifted for every source of the array
    want = o.model()
    rec = p.model()
    err = rel_l2(rec, want)
    print(f"  record rel_l2={err:.2e}", flush=True)
    assert np.abs(want).max() > 0 and err < 1e-4
    img, _ = o.rtm(want)
    ierr = rel_l2(p.rtm(want)["image"], img)
    print(f"  image rel_l2={ierr:.2e}", flush=True)
    assert ierr < 1e-4
