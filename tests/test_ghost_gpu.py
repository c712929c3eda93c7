"""Ghost mirrors on the GPU (Receiver::apply with ghost, Receiver::extract_value with mirror_ghost,
bgprtm/Receiver.cpp:394-426, 457-507) against the oracle, and whole shots with GHOST surfaces."""
import numpy as np
import pytest

from oracle import pipeline
from konglong_amd._lib import DeviceArray, KlrtmError, lib, ENGINE_FD, ISO, SNAP_FP32
from konglong_amd.propagator import Propagator, shot_config
from konglong_amd.receiver import Receiver
from tests.helpers import rel_l2, smooth_field

pytestmark = pytest.mark.gpu

SHAPE = (14, 17, 44)


def test_ghost_inject_extract_match_oracle(klrtm):
    ny, nx, nz = SHAPE
    izm = 12
    rng = np.random.default_rng(21)
    xyz = np.array([(x + 0.3, y + 0.6, izm + 2.5) for y in range(3, ny - 3, 2) for x in range(3, nx - 3, 2)],
                   np.float32)
    nr, nt = xyz.shape[0], 5
    print(f"[ghost] nr={nr} grid={SHAPE} izm={izm}", flush=True)
    traces = rng.standard_normal((nr, nt)).astype(np.float32)
    vel = smooth_field(SHAPE, 0.5, 1.5, 22)
    dt_dxyz = np.float32(1e-3)
    r = Receiver(nx, ny, nz, xyz, 5, 0.3, 0.3, nt, dt_dxyz)
    r.set_traces(traces)
    dw, dv = DeviceArray(SHAPE), DeviceArray.from_host(vel)
    r.apply(dw, 2, dv, 1.0, iz_mirror=izm)
    t = pipeline.spread_tables(nx, ny, nz, xyz, 5, 0.3, 0.3)
    want = np.zeros(SHAPE, np.float32)
    pipeline.inject(want, vel, SHAPE, t, traces[:, 2] * dt_dxyz, iz_mirror=izm)
    assert np.abs(want[:, :, :izm]).max() > 0
    assert rel_l2(dw.download(), want) < 1e-6
    w = rng.standard_normal(SHAPE).astype(np.float32)
    r.extract(DeviceArray.from_host(w), 1, iz_mirror=izm)
    assert rel_l2(r.record()[:, 1], pipeline.extract(w, SHAPE, t, iz_mirror=izm)) < 1e-6
    # one trace: the reference's products, bit for bit, mirror included
    one = Receiver(nx, ny, nz, xyz[5:6], 5, 0.3, 0.3, nt, dt_dxyz)
    one.set_traces(traces[5:6])
    d1 = DeviceArray(SHAPE)
    one.apply(d1, 2, dv, 1.0, iz_mirror=izm)
    w1 = np.zeros(SHAPE, np.float32)
    pipeline.inject(w1, vel, SHAPE, pipeline.spread_tables(nx, ny, nz, xyz[5:6], 5, 0.3, 0.3),
                    traces[5:6, 2] * dt_dxyz, iz_mirror=izm)
    assert np.array_equal(d1.download().view(np.uint32), w1.view(np.uint32))


def test_ghost_rejects_plane_outside_grid(klrtm):
    ny, nx, nz = SHAPE
    r = Receiver(nx, ny, nz, np.array([[8.0, 7.0, 20.0]], np.float32), 3, 0.3, 0.3, 2, 1e-3)
    d = DeviceArray(SHAPE)
    assert lib().klrtm_receiver_apply_ghost(r.h, d.ptr, 0, d.ptr, 1.0, nz) != 0
    assert lib().klrtm_receiver_extract_ghost(r.h, d.ptr, 0, -2) != 0


@pytest.mark.parametrize("sides", [(1, 1), (1, 0), (0, 1)], ids=["GHOST", "SOURCE_GHOST", "RECEIVER_GHOST"])
def test_ghost_shot_matches_oracle(klrtm, sides):
    print(f"[ghost-shot] sides={sides}", flush=True)
    ny, nx, nz = 34, 36, 44
    d, dt, nt, nb, izm = 10.0, 1e-3, 130, 10, 13
    vel = smooth_field((ny, nx, nz), 1800.0, 2600.0, 51)
    rcv = np.array([[x, y, izm + 2.0] for y in range(2, ny - 2, 3) for x in range(2, nx - 2, 3)], np.float32)
    src = (17.0, 16.0, izm + 3.0)
    cfg = shot_config(nx, ny, nz, d, dt, nt, ENGINE_FD, 8, ISO, nbnd=nb, vmax=2600.0, src_xyz=src, f_peak=15.0,
                      spread=6, khpass=0.3, khcut=0.3, nrcv=rcv.shape[0], snap_every=2, snap_format=SNAP_FP32,
                      tmute_frac=0.1)
    p = Propagator(cfg)
    p.set_model(vel)
    p.set_receivers(rcv)
    p.set_ghost(izm, *sides)
    o = pipeline.Shot(nx, ny, nz, d, dt, nt, ENGINE_FD, 8, ISO, [nb] * 6, 2600.0, src, 15.0, 6, 0.3, 0.3, rcv,
                      snap_every=2, tmute_frac=0.1, ghost=(izm, bool(sides[0]), bool(sides[1])))
    o.set_model(vel)
    p.set_source_wavelet(o.wavelet)
    want = o.model()
    rec = p.model()
    err = rel_l2(rec, want)
    print(f"  record rel_l2={err:.2e}", flush=True)
    assert np.abs(want).max() > 0 and err < 1e-4
    img_want, _ = o.rtm(want)
    ierr = rel_l2(p.rtm(want)["image"], img_want)
    print(f"  image rel_l2={ierr:.2e}", flush=True)
    assert np.abs(img_want).max() > 0 and ierr < 1e-4


def test_ghost_plane_must_fit(klrtm):
    cfg = shot_config(16, 16, 20, 10.0, 1e-3, 4, nbnd=4, vmax=2000.0, nrcv=1, spread=2)
    p = Propagator(cfg)
    with pytest.raises(KlrtmError):
        p.set_ghost(10)
    p.set_ghost(9)
    p.set_ghost(-1, False, False)
