"""Snapshot codecs on the GPU: WaveFieldCompress COMPRESS_16 and block-scaled FP16/BF16."""
import numpy as np

from ._lib import lib, check, DeviceArray, SNAP_FP16, SNAP_BF16  # noqa: F401


def nshort_volume(nz, nx, ny):
    return int(lib().klrtm_nshort_volume(nz, nx, ny))


class WaveFieldCompress:
    """COMPRESS_16 codec (bgprtm/WaveFieldCompress.h:4-53) on device arrays."""
    COMPRESS_16 = 1

    @staticmethod
    def nshort_volume(nz, nx, ny):
        return nshort_volume(nz, nx, ny)

    def compress(self, d_in, d_out, nx, ny, nz):
        check(lib().klrtm_compress16(d_in.ptr, d_out.ptr, nx, ny, nz))

    def uncompress(self, d_in, d_out, nx, ny, nz):
        check(lib().klrtm_uncompress16(d_in.ptr, d_out.ptr, nx, ny, nz))


def blockq_nscales(nx, ny, nz):
    return int(lib().klrtm_blockq_nscales(nx, ny, nz))


def blockq_encode(d_in, d_q, d_scales, nx, ny, nz, fmt):
    check(lib().klrtm_blockq_encode(d_in.ptr, d_q.ptr, d_scales.ptr, nx, ny, nz, fmt))


def blockq_decode(d_q, d_scales, d_out, nx, ny, nz, fmt):
    check(lib().klrtm_blockq_decode(d_q.ptr, d_scales.ptr, d_out.ptr, nx, ny, nz, fmt))
