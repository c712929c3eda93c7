"""The FFT imaging conditions on the GPU (konglong_amd/csrc/correlator.hip) against the FP64 restatement
(oracle/correlator.py) on band-limited fields: unit derivatives per axis, kfilter, costerm/sinterm and
every ImagingCorrelator::run option, in 3D and 2D.  Tolerance: 2e-6 relative L2 (FP32 FFTs)."""
import numpy as np
import pytest

from oracle import correlator as oc
from konglong_amd._lib import DeviceArray
from konglong_amd.imaging import ImagingCorrelator, IC_IMAGE1, IC_INVERSION, IC_INVERSION2, IC_ALLIN3
from tests.helpers import bandlimited_field, rel_l2

TOL = 2e-6
D = (10.0, 12.5, 7.5)


def _dev(a):
    return DeviceArray.from_host(np.asarray(a, np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(12, 18, 20), (1, 30, 24)])
def test_unit_derivatives_and_kfilter(klrtm, shape):
    ny, nx, nz = shape
    x = bandlimited_field(shape, 1).astype(np.float32)
    c = ImagingCorrelator(nx, ny, nz, D[0], D[1], D[2], kmin=0.02)
    out = DeviceArray(shape)
    for axis in ((0, 1, 2) if ny > 1 else (0, 1)):
        c.unit_derivative(_dev(x), axis, out)
        want = oc.unit_derivative(x, axis, D)
        assert rel_l2(out.download(), want) < TOL, axis
    w = _dev(x)
    c.kfilter(w)
    assert rel_l2(w.download(), oc.kfilter(x, D, 0.02)) < TOL
    img = _dev(x)
    c.costerm(img)
    assert rel_l2(img.download(), oc.costerm(x, D)) < TOL
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(12, 18, 20), (1, 30, 24)])
@pytest.mark.parametrize("option", [IC_IMAGE1, IC_INVERSION, IC_INVERSION2, IC_ALLIN3])
def test_run_options_match_oracle(klrtm, shape, option):
    ny, nx, nz = shape
    s = bandlimited_field(shape, 2).astype(np.float32)
    r = bandlimited_field(shape, 3).astype(np.float32)
    rng = np.random.default_rng(4)
    i0, f0, r0 = (rng.standard_normal(shape).astype(np.float32) * 0.1 for _ in range(3))
    c = ImagingCorrelator(nx, ny, nz, D[0], D[1], D[2], kmin=0.02)
    dimg, dfwi, draw = _dev(i0), _dev(f0), _dev(r0)
    d64 = DeviceArray(shape, np.float64)
    for step in range(2):   # two imaging steps accumulate
        c.run(_dev(s * (1 + step)), _dev(r), option, image=dimg, img_fwi=dfwi, img_raw=draw, image64=d64)
    wi, wf, wr = i0.astype(np.float64), f0.astype(np.float64), r0.astype(np.float64)
    w64 = np.zeros(shape)
    for step in range(2):
        wi, wf, wr = oc.run(option, s * (1 + step), r, D, 0.02, wi, wf, wr)
        w64 = oc.run(option, s * (1 + step), r, D, 0.02, w64, None, None)[0]
    assert rel_l2(dimg.download(), wi) < TOL
    assert rel_l2(d64.download(), w64) < TOL
    if option in (IC_INVERSION, IC_INVERSION2, IC_ALLIN3):
        assert rel_l2(dfwi.download(), wf) < TOL
    else:
        assert np.array_equal(dfwi.download(), f0)
    if option == IC_ALLIN3:
        assert rel_l2(draw.download(), wr) < TOL
    else:
        assert np.array_equal(draw.download(), r0)
    if option == IC_INVERSION2:   # the post-loop steps
        c.costerm(dimg)
        c.sinterm(dfwi, dimg)
        wi = oc.costerm(wi, D)
        assert rel_l2(dimg.download(), wi) < TOL
        assert rel_l2(dfwi.download(), oc.sinterm(wf, wi)) < TOL
    c.close()


@pytest.mark.gpu
def test_plane_wave_inversion_is_cos_2theta(klrtm):
    ny, nx, nz = 8, 10, 12
    y, x, z = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    th = 2 * np.pi * (1 * y / ny + 2 * x / nx + 3 * z / nz)
    cs = np.cos(th).astype(np.float32)
    c = ImagingCorrelator(nx, ny, nz, D[0], D[1], D[2])
    img, fwi = DeviceArray((ny, nx, nz)), DeviceArray((ny, nx, nz))
    c.run(_dev(cs), _dev(cs), IC_INVERSION, image=img, img_fwi=fwi)
    assert np.abs(img.download() - np.cos(2 * th)).max() < 1e-5
    assert np.abs(fwi.download() - 1.0).max() < 1e-5
    c.close()


def test_odd_nz_rejected():
    from konglong_amd._lib import lib
    import ctypes as C
    h = C.c_void_p()
    assert lib().klrtm_correlator_create(C.byref(h), 8, 8, 9, 1.0, 1.0, 1.0, 0.0) == 1
    assert b"even" in lib().klrtm_last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("option", [IC_INVERSION, IC_INVERSION2, IC_ALLIN3], ids=["inversion", "inversion2", "allin3"])
def test_rtm_with_fft_imaging_matches_oracle(klrtm, option):
    """A whole FD8 ISO shot imaged with an FFT imaging condition: image, FWI, raw and illumination against the
    oracle pipeline (FP32 shot, FP64 imaging condition), 1e-4 relative L2."""
    from konglong_amd._lib import ENGINE_FD, ISO
    from tests.test_pipeline_gpu import _shot
    print(f"[rtm-ic] option {option}", flush=True)
    p, o = _shot(ENGINE_FD, 8, ISO)
    kmin = 1.0 / 1492.0
    p.set_imaging(option, kmin)
    rec = o.model()
    out = p.rtm(rec, want=("image", "illum"))
    img, ill, fwi, raw = o.rtm(rec, ic_option=option, kmin=kmin)
    errs = [rel_l2(out["image"], img), rel_l2(out["illum"], ill), rel_l2(out["fwi"], fwi)]
    if option == IC_ALLIN3:
        errs.append(rel_l2(out["raw"], raw))
    print(f"  rel_l2 image/illum/fwi(/raw) = {errs}", flush=True)
    assert np.abs(img).max() > 0 and np.abs(fwi).max() > 0
    assert max(errs) < 1e-4
    p.close()
