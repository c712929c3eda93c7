"""Imaging on a doubled grid (ic_double_z/x/y) on the GPU against the FP64 restatement (oracle/correlator.py):
the high cut and half-sample copies (relative L2 below 1e-5: FP32 line FFTs), the interleave and the
illumination average (bit-exact), ImagingCorrelator::run for IMAGE1 and INVERSION, and a whole FD8 ISO shot
through the propagator (relative L2 below 1e-4, as the other FFT imaging conditions)."""
import ctypes as C

import numpy as np
import pytest

from oracle import correlator as oc
from konglong_amd._lib import DeviceArray, lib, KlrtmError
from konglong_amd.imaging import ImagingCorrelator, IC_IMAGE1, IC_INVERSION, IC_INVERSION2
from tests.helpers import bandlimited_field, rel_l2

TOL = 1e-5
D = (10.0, 12.5, 7.5)
LAYOUTS = [((2, 2, 2), (12, 18, 20)), ((2, 2, 1), (1, 30, 24)), ((2, 1, 1), (1, 30, 24))]
IDS = ["3d-zxy", "2d-zx", "2d-z"]


def _dev(a):
    return DeviceArray.from_host(np.asarray(a, np.float32))


def _corr(dbl, shape, high=(0.9, 0.7)):
    ny, nx, nz = shape
    c = ImagingCorrelator(nx, ny, nz, D[0], D[1], D[2], kmin=0.02)
    c.set_double(*dbl, *high)
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("dbl,shape", LAYOUTS, ids=IDS)
def test_double_pieces_match_oracle(klrtm, dbl, shape):
    print(f"[icdouble pieces] {dbl} {shape}", flush=True)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(shape).astype(np.float32)   # full band: the high cut acts
    c = _corr(dbl, shape)
    nl = c.nloop
    assert nl == len(oc.double_copies(*dbl))
    w = _dev(x)
    c.hightape(w)
    assert rel_l2(w.download(), oc.hightape(x, dbl)) < TOL
    buf = DeviceArray((nl,) + shape)
    y = bandlimited_field(shape, 12, cut=0.8).astype(np.float32)
    buf.upload(np.concatenate([y[None], np.zeros((nl - 1,) + shape, np.float32)]))
    c.interp(buf)
    got = buf.download()
    for i, want in enumerate(oc.interp(y, dbl)):
        assert rel_l2(got[i], want) < TOL, i
    sub = rng.standard_normal((nl,) + shape).astype(np.float32)
    dz, dx, dy = dbl
    big = (shape[0] * dy, shape[1] * dx, shape[2] * dz)
    out = DeviceArray(big)
    c.interleave(_dev(sub), out)
    assert np.array_equal(out.download(), oc.interleave(sub, dbl))
    ill = rng.random(shape).astype(np.float32)
    c.interp_illum(_dev(ill), out)
    assert np.array_equal(out.download(), oc.interp_illum(ill, dbl))
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dbl,shape", LAYOUTS, ids=IDS)
@pytest.mark.parametrize("option", [IC_IMAGE1, IC_INVERSION], ids=["image1", "inversion"])
def test_double_run_matches_oracle(klrtm, dbl, shape, option):
    print(f"[icdouble run] {dbl} {shape} option {option}", flush=True)
    s = bandlimited_field(shape, 2, cut=0.95).astype(np.float32)
    r = bandlimited_field(shape, 3, cut=0.95).astype(np.float32)
    c = _corr(dbl, shape)
    nl = c.nloop
    img, fwi = DeviceArray((nl,) + shape), DeviceArray((nl,) + shape)
    img.upload(np.zeros((nl,) + shape, np.float32))
    fwi.upload(np.zeros((nl,) + shape, np.float32))
    wi, wf = np.zeros((nl,) + shape), np.zeros((nl,) + shape)
    for step in range(2):   # two imaging steps accumulate
        ds = _dev(s * (1 + step))
        c.run(ds, _dev(r), option, image=img, img_fwi=fwi)
        sh, wi, wf = oc.run_double(option, s * (1 + step), r, D, dbl, (0.9, 0.7), wi, wf)
        assert rel_l2(ds.download(), sh) < TOL   # src is high-cut in place, as applyFFTV leaves srcWF[0]
    assert rel_l2(img.download(), wi) < TOL
    if option == IC_INVERSION:
        assert rel_l2(fwi.download(), wf) < TOL
    else:
        assert not np.any(fwi.download())
    with pytest.raises(KlrtmError):
        c.run(_dev(s), _dev(r), option, image=img, image64=DeviceArray(shape, np.float64))
    with pytest.raises(KlrtmError):
        c.run(_dev(s), _dev(r), IC_INVERSION2, image=img)
    c.close()


@pytest.mark.gpu
def test_double_layout_rules(klrtm):
    h = C.c_void_p()
    assert lib().klrtm_correlator_create(C.byref(h), 8, 6, 10, 1.0, 1.0, 1.0, 0.0) == 0
    for dz, dx, dy in ((2, 2, 1), (2, 1, 1), (1, 2, 2), (3, 1, 1)):   # 3D doubles all axes or none
        assert lib().klrtm_correlator_set_double(h, dz, dx, dy, 0.9, 0.7) != 0
    lib().klrtm_correlator_destroy(h)
    assert lib().klrtm_correlator_create(C.byref(h), 8, 1, 10, 1.0, 1.0, 1.0, 0.0) == 0
    assert lib().klrtm_correlator_set_double(h, 2, 2, 2, 0.9, 0.7) != 0   # no y on a 2D grid
    assert lib().klrtm_correlator_set_double(h, 1, 2, 1, 0.9, 0.7) != 0   # x needs z
    assert lib().klrtm_correlator_set_double(h, 2, 2, 1, 0.9, 0.7) == 0
    lib().klrtm_correlator_destroy(h)


@pytest.mark.gpu
@pytest.mark.parametrize("option", [IC_IMAGE1, IC_INVERSION], ids=["image1", "inversion"])
def test_rtm_on_doubled_grid_matches_oracle(klrtm, option):
    """A whole FD8 ISO shot imaged on the doubled grid: image, illumination (and the FWI image) against the
    oracle pipeline (FP32 shot, FP64 imaging condition), 1e-4 relative L2."""
    from konglong_amd._lib import ENGINE_FD, ISO
    from tests.test_pipeline_gpu import _shot
    print(f"[rtm-icdouble] option {option}", flush=True)
    p, o = _shot(ENGINE_FD, 8, ISO)
    if option != IC_IMAGE1:
        p.set_imaging(option)
    p.set_ic_double(2, 2, 2)
    rec = o.model()
    out = p.rtm(rec)
    img, ill, fwi = o.rtm(rec, ic_option=option, dbl=(2, 2, 2))
    errs = [rel_l2(out["image"], img), rel_l2(out["illum"], ill)]
    if option == IC_INVERSION:
        errs.append(rel_l2(out["fwi"], fwi))
    print(f"  rel_l2 image/illum(/fwi) = {errs}", flush=True)
    assert out["image"].shape == (80, 84, 88) and np.abs(img).max() > 0
    assert max(errs) < 1e-4
    from konglong_amd._lib import fptr   # host copies of the propagation-grid size are refused
    assert lib().klrtm_propagator_rtm(p.h, None, fptr(np.zeros(p.shape, np.float32)), None, None, None) != 0
    p.set_ic_double(1, 1, 1)
    plain = p.rtm(rec, want=("image", "illum"))
    assert plain["image"].shape == p.shape
    p.set_imaging(IC_INVERSION2)   # refused on a doubled grid at set-up, whichever comes first
    with pytest.raises(KlrtmError):
        p.set_ic_double(2, 2, 2)
    p.set_imaging(option)
    p.set_ic_double(2, 2, 2)
    with pytest.raises(KlrtmError):
        p.set_imaging(IC_INVERSION2)
    p.close()


@pytest.mark.gpu
@pytest.mark.parametrize("option", [IC_IMAGE1, IC_INVERSION], ids=["image1", "inversion"])
def test_run_copies_is_the_reference_contract(klrtm, option):
    """run() = applyFFTV (+ interp for IMAGE1, as stackImageInterp does) followed by run_copies on the
    reference's buffers: the same kernels in the same order, so the images agree bit for bit."""
    dbl, shape = (2, 2, 2), (10, 12, 16)
    s = bandlimited_field(shape, 7, cut=0.95).astype(np.float32)
    r = bandlimited_field(shape, 8, cut=0.95).astype(np.float32)
    c = _corr(dbl, shape)
    nl = c.nloop
    a_img, a_fwi, b_img, b_fwi = (DeviceArray((nl,) + shape) for _ in range(4))
    for d in (a_img, a_fwi, b_img, b_fwi):
        d.upload(np.zeros((nl,) + shape, np.float32))
    c.run(_dev(s), _dev(r), option, image=a_img, img_fwi=a_fwi)
    sb, rb = DeviceArray((nl,) + shape), DeviceArray((nl,) + shape)
    sb.upload(np.concatenate([s[None], np.zeros((nl - 1,) + shape, np.float32)]))
    rb.upload(np.concatenate([r[None], np.zeros((nl - 1,) + shape, np.float32)]))
    c.hightape(sb)   # copy 0
    c.hightape(rb)
    if option == IC_IMAGE1:
        c.interp(sb)
        c.interp(rb)
    c.run_copies(sb, rb, option, image=b_img, img_fwi=b_fwi)
    assert np.array_equal(a_img.download(), b_img.download())
    assert np.array_equal(a_fwi.download(), b_fwi.download())
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("option", [IC_IMAGE1, IC_INVERSION], ids=["image1", "inversion"])
def test_rtm_on_doubled_grid_2d(klrtm, option):
    """A 2D FD8 ISO shot imaged with z and x doubled (interleave2d layout) against the oracle pipeline."""
    from konglong_amd._lib import ENGINE_FD, ISO
    from konglong_amd.propagator import Propagator, shot_config
    from oracle import pipeline
    from tests.helpers import smooth_field
    print(f"[rtm-icdouble-2d] option {option}", flush=True)
    ny, nx, nz, d, dt, nt = 1, 72, 56, 10.0, 1e-3, 180
    nb = [10, 10, 0, 0, 10, 10]
    vel = smooth_field((ny, nx, nz), 1800.0, 2600.0, 31)
    rcv = np.array([[x, 0.0, 12.0] for x in range(3, nx - 3, 2)], np.float32)
    src = (35.0, 0.0, 14.0)
    cfg = shot_config(nx, ny, nz, d, dt, nt, ENGINE_FD, 8, ISO, nbnd=nb, vmax=2600.0, src_xyz=src, f_peak=15.0,
                      spread=6, nrcv=rcv.shape[0], snap_every=2, tmute_frac=0.1)
    p = Propagator(cfg)
    p.set_model(vel)
    p.set_receivers(rcv)
    o = pipeline.Shot(nx, ny, nz, d, dt, nt, ENGINE_FD, 8, ISO, nb, 2600.0, src, 15.0, 6, 0.3, 0.3, rcv,
                      snap_every=2, tmute_frac=0.1)
    o.set_model(vel)
    p.set_source_wavelet(o.wavelet)
    if option != IC_IMAGE1:
        p.set_imaging(option)
    p.set_ic_double(2, 2, 1)
    rec = o.model()
    out = p.rtm(rec)
    img, ill, fwi = o.rtm(rec, ic_option=option, dbl=(2, 2, 1))
    errs = [rel_l2(out["image"], img), rel_l2(out["illum"], ill)]
    if option == IC_INVERSION:
        errs.append(rel_l2(out["fwi"], fwi))
    print(f"  rel_l2 image/illum(/fwi) = {errs}", flush=True)
    assert out["image"].shape == (1, 2 * nx, 2 * nz) and np.abs(img).max() > 0
    assert max(errs) < 1e-4
    p.close()
