"""Q attenuation on the GPU: the Q branches of Propagator::Vel2 and Propagator::update2nd
(bgprtm/Propagator.cpp:725-741, 1789-1816) bit for bit against the oracle, and whole shots with a
1/Q model (kernel, then the Q leapfrog with swapping memory-variable sets) against the oracle shot."""
import numpy as np
import pytest

import oracle
from oracle import pipeline
from konglong_amd._lib import DeviceArray, KlrtmError, check, lib, fptr, ENGINE_FD, ENGINE_CFD, ISO, VTI, SNAP_FP32
from konglong_amd.propagator import Propagator, shot_config
from tests.helpers import rel_l2, smooth_field

pytestmark = pytest.mark.gpu


def _coeffs(order):
    return np.array([0.04] + [0.3, 0.2, 0.1, 0.05, 0.02, 0.01, 0.005, 0.002][:order] +
                    [0.9, 0.7, 0.5, 0.3, 0.1, 0.05, 0.02, 0.01][:order], np.float32)


@pytest.mark.parametrize("order", [1, 4, 8])
def test_update2nd_q_bit_exact(klrtm, order):
    print(f"[q-update] order={order}", flush=True)
    shape = (5, 13, 37)
    rng = np.random.default_rng(order)
    w0 = rng.standard_normal(shape).astype(np.float32)
    w1 = rng.standard_normal(shape).astype(np.float32)
    iq = rng.uniform(0, 0.1, shape).astype(np.float32)
    sets = [rng.standard_normal((order,) + shape).astype(np.float32) for _ in range(2)]
    c = _coeffs(order)
    dw0, dw1, diq = DeviceArray.from_host(w0), DeviceArray.from_host(w1), DeviceArray.from_host(iq)
    dsets = [DeviceArray.from_host(s) for s in sets]
    n = w0.size
    # three steps with the sets swapping, as the propagator runs them
    for step in range(3):
        oracle.update2nd_q(w0, w1, iq, sets[0], sets[1], c)
        check(lib().klrtm_update2nd_q(dw0.ptr, dw1.ptr, diq.ptr, dsets[0].ptr, dsets[1].ptr, n, order, fptr(c)))
        w0, w1 = w1, w0
        dw0, dw1 = dw1, dw0
        sets.reverse()
        dsets.reverse()
    assert np.array_equal(dw1.download().view(np.uint32), w1.view(np.uint32))
    for s, d in zip(sets, dsets):
        assert np.array_equal(d.download().view(np.uint32), s.view(np.uint32))


def test_vel2_q_bit_exact(klrtm):
    shape = (4, 9, 21)
    rng = np.random.default_rng(5)
    v = rng.uniform(1500, 4500, shape).astype(np.float32)
    iq = rng.uniform(0, 0.1, shape).astype(np.float32)
    c = _coeffs(3)
    dv, diq = DeviceArray.from_host(v), DeviceArray.from_host(iq)
    check(lib().klrtm_vel2_q(dv.ptr, diq.ptr, v.size, 7e-4, float(c[0])))
    oracle.vel2_q(v, iq, np.float32(7e-4), c)
    assert np.array_equal(dv.download().view(np.uint32), v.view(np.uint32))
    assert np.array_equal(diq.download().view(np.uint32), iq.view(np.uint32))


def test_update2nd_q_rejects_bad_order(klrtm):
    d = DeviceArray((64,))
    c = _coeffs(8)
    assert lib().klrtm_update2nd_q(d.ptr, d.ptr, d.ptr, d.ptr, d.ptr, 64, 0, fptr(c)) != 0
    assert lib().klrtm_update2nd_q(d.ptr, d.ptr, d.ptr, d.ptr, d.ptr, 64, 9, fptr(c)) != 0


@pytest.mark.parametrize("case", [(ENGINE_FD, 8, ISO), (ENGINE_CFD, 10, VTI), (ENGINE_FD, 8, VTI)],
                         ids=["FD8-ISO", "CFD10-VTI", "FD8-VTI"])
def test_q_shot_matches_oracle(klrtm, case):
    engine, order, model = case
    print(f"[q-shot] {case}", flush=True)
    ny, nx, nz = 34, 36, 40
    d, dt, nt, nb = 10.0, 1e-3, 130, 10
    vel = smooth_field((ny, nx, nz), 1800.0, 2600.0, 41)
    eps = smooth_field((ny, nx, nz), 0.0, 0.2, 42) if model == VTI else None
    dl = smooth_field((ny, nx, nz), 0.0, 0.1, 43) if model == VTI else None
    invq = smooth_field((ny, nx, nz), 0.01, 0.08, 44)
    coeffs = _coeffs(3)
    rcv = np.array([[x, y, nb + 2.0] for y in range(2, ny - 2, 3) for x in range(2, nx - 2, 3)], np.float32)
    src = (17.0, 16.0, nb + 4.0)
    k1, k2 = (0.3, 0.3) if engine == ENGINE_FD else (0.40, 0.48)
    cfg = shot_config(nx, ny, nz, d, dt, nt, engine, order, model, nbnd=nb, vmax=2600.0, src_xyz=src, f_peak=15.0,
                      spread=6, khpass=k1, khcut=k2, nrcv=rcv.shape[0], snap_every=2, snap_format=SNAP_FP32,
                      tmute_frac=0.1)
    p = Propagator(cfg)
    p.set_q(invq, coeffs)
    p.set_model(vel, eps, dl)
    p.set_receivers(rcv)
    o = pipeline.Shot(nx, ny, nz, d, dt, nt, engine, order, model, [nb] * 6, 2600.0, src, 15.0, 6, k1, k2, rcv,
                      snap_every=2, tmute_frac=0.1, q=(invq, coeffs))
    o.set_model(vel, eps, dl)
    p.set_source_wavelet(o.wavelet)
    want = o.model()
    rec = p.model()
    err = rel_l2(rec, want)
    print(f"  record rel_l2={err:.2e}", flush=True)
    assert np.abs(want).max() > 0 and err < 1e-4
    img_want, _ = o.rtm(want)
    got = p.rtm(want)
    ierr = rel_l2(got["image"], img_want)
    print(f"  image rel_l2={ierr:.2e}", flush=True)
    assert np.abs(img_want).max() > 0 and ierr < 1e-4
    # Q changed the shot: the same propagator with Q switched off records something else
    p.set_q(None, None)
    p.set_model(vel, eps, dl)
    assert rel_l2(p.model(), want) > 1e-3


def test_q_needs_set_model_after_set_q(klrtm):
    ny, nx, nz = 16, 16, 20
    cfg = shot_config(nx, ny, nz, 10.0, 1e-3, 4, nbnd=4, vmax=2000.0, nrcv=1, spread=2)
    p = Propagator(cfg)
    vel = np.full((ny, nx, nz), 2000.0, np.float32)
    p.set_model(vel)
    p.set_receivers(np.array([[8.0, 8.0, 6.0]], np.float32))
    p.set_q(np.full((ny, nx, nz), 0.02, np.float32), _coeffs(2))
    with pytest.raises(KlrtmError, match="set_model"):
        p.model()
    p.set_model(vel)
    p.model()
    # switching Q off also needs set_model: the velocity still carries the Q correction of Vel2
    p.set_q(None, None)
    with pytest.raises(KlrtmError, match="set_model"):
        p.model()
    p.set_model(vel)
    p.model()
