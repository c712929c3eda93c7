"""Illumination compensation on the GPU against the oracle: the box smoothing (avgVolume3D) bit for
bit, the damping weights to FP32 rounding (exp on the device), and the image weighting."""
import numpy as np
import pytest

import oracle
from konglong_amd._lib import DeviceArray, check, lib
from tests.helpers import rel_l2, smooth_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,h", [((9, 30, 11), (3, 2, 1)), ((24, 26, 40), (10, 10, 10)), ((5, 7, 64), (0, 4, 6))])
def test_avg_volume3d_bit_exact(klrtm, shape, h):
    print(f"[postproc] box {shape} {h}", flush=True)
    ny, nx, nz = shape
    v = np.random.default_rng(sum(shape)).standard_normal(shape).astype(np.float32)
    want = oracle.avg_volume3d(v.copy(), *h)
    d = DeviceArray.from_host(v)
    check(lib().klrtm_avg_volume3d(d.ptr, nx, ny, nz, *h))
    assert np.array_equal(d.download().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("smooth", [1, 0])
def test_illum_damping_matches_oracle(klrtm, smooth):
    import ctypes as C
    shape = (30, 34, 28)
    ny, nx, nz = shape
    ill = (smooth_field(shape, 0.1, 5.0, 41) ** 2).astype(np.float32)
    want = ill.copy()
    avg = oracle.illum_damping(want, smooth=bool(smooth), smooth_xy=4, smooth_z=3, amp_damping=0.9, damp_weight=0.07)
    d = DeviceArray.from_host(ill)
    a = C.c_double()
    check(lib().klrtm_illum_damping(d.ptr, nx, ny, nz, smooth, 4, 3, 0.9, 0.07, C.byref(a)))
    got = d.download()
    print(f"[postproc] damping smooth={smooth}: average {a.value:.9g} vs {avg:.9g}, rel_l2 {rel_l2(got, want):.2e}",
          flush=True)
    assert abs(a.value - avg) <= 1e-12 * abs(avg)
    assert rel_l2(got, want) < 1e-6 and np.allclose(got, want, rtol=1e-6, atol=0)
    img = smooth_field(shape, -1.0, 1.0, 42)
    di = DeviceArray.from_host(img)
    check(lib().klrtm_apply_damping(d.ptr, di.ptr, img.size))
    assert np.array_equal(di.download(), (img * got).astype(np.float32))
