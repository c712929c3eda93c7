"""The bgprtm drop-in C++ headers, driven from a compiled program (tests/cpp/api_check.cpp)
against the oracle: KernelCPU::VTI (CFD, bit-exact), FdEngine_fd::laplacian,
WaveFieldCompress COMPRESS_16P, PML::apply2_single_trans, ImagingCorrelator::crossone."""
import subprocess

import pytest

from tests.helpers import build_api_check

pytestmark = pytest.mark.gpu


def test_cpp_api_check(klrtm, tmp_path):
    exe = build_api_check(tmp_path / "api_check")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "API_CHECK PASS" in r.stdout, r.stdout + r.stderr
