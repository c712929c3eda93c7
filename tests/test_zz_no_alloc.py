"""The allocation-free COMPRESS_16 entry point (klrtm_compress16_noalloc) against klrtm_compress16.

Both are called through the C ABI on device buffers; the codes must agree bit for bit on two calls
in a row, the second with a smaller volume (the scratch of slice maxima keeps the first call's
contents, so the second call proves it is cleared), and the oracle's COMPRESS_16 must agree too."""
import ctypes as C

import numpy as np
import pytest

import oracle
from konglong_amd._lib import DeviceArray
from konglong_amd.codec import nshort_volume

MAX_NY = 32768  # KLRTM_C16_NOALLOC_MAX_NY in include/konglong_rtm.h


def _encode(L, fn, v):
    ny, nx, nz = v.shape
    dv = DeviceArray.from_host(v)
    dc = DeviceArray((nshort_volume(nz, nx, ny),), np.uint16)
    assert getattr(L, fn)(dv.ptr, dc.ptr, nx, ny, nz) == 0, L.klrtm_last_error()
    assert L.klrtm_synchronize() == 0
    return dc.download()


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point(klrtm):
    L = klrtm.lib()
    rng = np.random.default_rng(11)
    for shape in ((40, 44, 36), (9, 21, 30)):   # second call smaller, with larger slice maxima
        v = (rng.standard_normal(shape) * np.exp(rng.uniform(-6, 6, shape))).astype(np.float32)
        if shape[0] == 9:
            v *= np.float32(1e-3)               # maxima below the first call's: stale maxima would show
            v[3] = 0.0                          # an all-zero slice
        got = _encode(L, "klrtm_compress16_noalloc", v)
        want = _encode(L, "klrtm_compress16", v)
        assert np.array_equal(got, want)
        assert np.array_equal(got, oracle.compress16(v))


@pytest.mark.gpu
def test_noalloc_rejects_oversized_volume(klrtm):
    L = klrtm.lib()
    dv = DeviceArray((4,))
    assert L.klrtm_compress16_noalloc(dv.ptr, dv.ptr, 1, MAX_NY + 1, 1) == 1
    assert b"exceeds" in L.klrtm_last_error()
    assert L.klrtm_compress16_noalloc(None, dv.ptr, 1, 1, 1) == 1
