"""Free surface on the GPU: Propagator::apply_symmetry bit for bit against the oracle, and whole
shots with a FREESURFACE top (top PML off, mirror after every injection) against the oracle shot."""
import numpy as np
import pytest

import oracle
from oracle import pipeline
from konglong_amd._lib import DeviceArray, check, lib, ENGINE_FD, ENGINE_CFD, ISO, VTI, SNAP_FP32
from konglong_amd.propagator import Propagator, shot_config
from tests.helpers import rel_l2, smooth_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("iz0,tpr,sym,taper", [(9, 6, -1, 1), (9, 9, -1, 1), (12, 5, 1, 1), (7, 0, -1, 0),
                                               (10, 0, 1, 0), (3, 3, -1, 1)])
def test_apply_symmetry_bit_exact(klrtm, iz0, tpr, sym, taper):
    ny, nx, nz = 7, 33, 29
    print(f"[surface] iz0={iz0} tprlen={tpr} sym={sym} taper={taper}", flush=True)
    v = np.random.default_rng(iz0 + tpr).standard_normal((ny, nx, nz)).astype(np.float32)
    want = oracle.apply_symmetry(v.copy(), iz0, sym, tpr, taper)
    d = DeviceArray.from_host(v)
    check(lib().klrtm_apply_symmetry(d.ptr, nx, ny, nz, iz0, sym, tpr, taper))
    assert np.array_equal(d.download().view(np.uint32), want.view(np.uint32))


def test_apply_symmetry_rejects_bad_arguments(klrtm):
    d = DeviceArray((2, 3, 20))
    assert lib().klrtm_apply_symmetry(d.ptr, 3, 2, 20, 10, -1, 4, 1) != 0     # 2*iz0 + 2 >= nz
    assert lib().klrtm_apply_symmetry(d.ptr, 3, 2, 20, 5, -1, 6, 1) != 0      # taper longer than iz0
    assert lib().klrtm_apply_symmetry(d.ptr, 3, 2, 20, 5, 2, 4, 1) != 0       # sym not +-1


@pytest.mark.parametrize("case", [(ENGINE_FD, 8, ISO), (ENGINE_CFD, 10, VTI)], ids=["FD8-ISO", "CFD10-VTI"])
def test_free_surface_shot_matches_oracle(klrtm, case):
    engine, order, model = case
    print(f"[surface-shot] {case}", flush=True)
    ny, nx, nz = 36, 38, 44
    d, dt, nt, nb, iz0 = 10.0, 1e-3, 140, 10, 12
    vel = smooth_field((ny, nx, nz), 1800.0, 2600.0, 31)
    eps = smooth_field((ny, nx, nz), 0.0, 0.2, 32) if model == VTI else None
    dl = smooth_field((ny, nx, nz), 0.0, 0.1, 33) if model == VTI else None
    rcv = np.array([[x, y, iz0 + 2.0] for y in range(2, ny - 2, 3) for x in range(2, nx - 2, 3)], np.float32)
    src = (19.0, 17.0, iz0 + 3.0)
    k1, k2 = (0.3, 0.3) if engine == ENGINE_FD else (0.40, 0.48)
    cfg = shot_config(nx, ny, nz, d, dt, nt, engine, order, model, nbnd=nb, vmax=2600.0, src_xyz=src, f_peak=15.0,
                      spread=6, khpass=k1, khcut=k2, nrcv=rcv.shape[0], snap_every=2, snap_format=SNAP_FP32,
                      tmute_frac=0.1, free_surface_iz=iz0, surface_taper=1)
    p = Propagator(cfg)
    p.set_model(vel, eps, dl)
    p.set_receivers(rcv)
    o = pipeline.Shot(nx, ny, nz, d, dt, nt, engine, order, model, [nb] * 6, 2600.0, src, 15.0, 6, k1, k2, rcv,
                      snap_every=2, tmute_frac=0.1, free_surface_iz=iz0, surface_taper=1)
    o.set_model(vel, eps, dl)
    p.set_source_wavelet(o.wavelet)
    want = o.model()
    rec = p.model()
    err = rel_l2(rec, want)
    print(f"  record rel_l2={err:.2e}", flush=True)
    assert np.abs(want).max() > 0 and err < 1e-4
    if engine == ENGINE_CFD:
        assert np.array_equal(p.wavefield().view(np.uint32), o.last.view(np.uint32))
    out = p.rtm(want)
    img, _ = o.rtm(want)
    e = rel_l2(out["image"], img)
    print(f"  image rel_l2={e:.2e}", flush=True)
    assert e < 1e-4
    p.close()
