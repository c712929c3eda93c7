"""C ABI checks that need no GPU: the library loads, exports every entry point
the header declares, host-side set-up matches the reference / oracle bit for
bit, and invalid arguments come back as error codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from konglong_amd import _lib
from konglong_amd.receiver import spread_coeff

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    syms = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            syms |= set(re.findall(r"\b(klrtm_[a-z0-9_]+)\s*\(", txt))
    return syms


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = _header_symbols()
    assert len(syms) > 40
    missing = [s for s in sorted(syms) if not hasattr(L, s)]
    assert not missing, missing
    assert set(_lib.declared_symbols()) <= syms


@pytest.mark.parametrize("n,d", [(10, 1.0), (128, 10.0), (400, 12.5), (61, 7.0)])
def test_cfd_plan_bit_exact_with_reference(n, d):
    P = C.POINTER(C.c_float)
    bm, bl, am = np.zeros(10 * n, np.float32), np.zeros(8 * n, np.float32), np.zeros(120, np.float32)
    _lib.check(_lib.lib().klrtm_cfd_plan(n, d, bm.ctypes.data_as(P), bl.ctypes.data_as(P), am.ctypes.data_as(P)))
    ref = oracle.ref_cfd_lib()
    want = oracle.cfd_plan(n, d) if ref is None else None
    if ref is not None:
        want = [np.zeros(10 * n, np.float32), np.zeros(8 * n, np.float32), np.zeros(120, np.float32)]
        ref.ref_cfd_plan(n, d, *[x.ctypes.data_as(P) for x in want])
    for x, y in zip((bm, bl, am), want):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("x", [0.0, 3.0, 20.0, 20.3, 64.5, 126.9, 127.0])
@pytest.mark.parametrize("band", [(0.3, 0.3), (0.40, 0.48), (0.3575, 0.3575)])
def test_spread_coeff_bit_exact_with_oracle(x, band):
    off, k = spread_coeff(128, x, band[0], band[1], 20)
    ooff, ok = oracle.spread_coeff(128, x, band[0], band[1], 20)
    assert off == ooff and np.array_equal(k.view(np.uint32), ok.view(np.uint32))


def test_invalid_arguments_return_error_codes():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.klrtm_kernel_create(C.byref(h), 2, 10, 10, 1.0, 1.0, 1.0, 7, 8) == 1
    assert b"nx" in L.klrtm_last_error()
    assert L.klrtm_kernel_create(C.byref(h), 10, 10, 10, 1.0, 1.0, 1.0, 3, 8) == 1
    assert L.klrtm_kernel_create(C.byref(h), 10, 10, 10, 1.0, 1.0, 1.0, 7, 18) == 1
    assert L.klrtm_kernel_create(C.byref(h), 9, 10, 10, 1.0, 1.0, 1.0, 4, 10) == 1
    P = C.POINTER(C.c_float)
    z = np.zeros(40, np.float32)
    assert L.klrtm_cfd_plan(5, 1.0, z.ctypes.data_as(P), z.ctypes.data_as(P), z.ctypes.data_as(P)) == 1
    nl, off = C.c_int(20), C.c_int(0)
    assert L.klrtm_spread_coeff(128, z.ctypes.data_as(P), 10.0, 0.4, 0.3, C.byref(nl), C.byref(off)) == 1
    assert L.klrtm_spread_coeff(128, z.ctypes.data_as(P), 200.0, 0.3, 0.3, C.byref(nl), C.byref(off)) == 1
    assert L.klrtm_compress16(None, None, 4, 4, 4) == 1
    assert L.klrtm_ic_correlate(None, None, 4, 4, 4, 1.0, None, 0, None, None, None, None) == 1


def test_nshort_volume_matches_reference_formula():
    # bgprtm/WaveFieldCompress.cpp:101-106: (nz*nx + 2) * ny ushorts
    assert _lib.lib().klrtm_nshort_volume(128, 128, 128) == (128 * 128 + 2) * 128
    assert _lib.lib().klrtm_blockq_nscales(400, 400, 400) == 50 ** 3
    assert _lib.lib().klrtm_blockq_nscales(9, 17, 8) == 2 * 3 * 1


def test_cpp_dropin_headers_build(tmp_path):
    """The bgprtm drop-in headers compile warning-free and link against the library (no GPU call)."""
    from tests.helpers import build_api_check
    exe = build_api_check(tmp_path / "api_check")
    assert os.path.exists(exe)

