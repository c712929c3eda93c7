"""Snapshot codecs on the GPU: WaveFieldCompress COMPRESS_16 and block-scaled FP16/BF16."""
import numpy as np

from ._lib import lib, check, DeviceArray, SNAP_FP16, SNAP_BF16  # noqa: F401


def nshort_volume(nz, nx, ny):
    return int(lib().klrtm_nshort_volume(nz, nx, ny))


class WaveFieldCompress:
    """WaveFieldCompress (bgprtm/WaveFieldCompress.h:4-53) on device arrays.

    ``compression`` selects the codec like global_pars "wf_compress": COMPRESS_16
    (default, per-y-slice 16-bit code), COMPRESS_16P / COMPRESS_8P (per-z-trace
    sum/difference codes) or COMPRESS_NONE (whole floats).  Decoding applies the
    reference's x/y/z edge tapers.
    """
    COMPRESS_NONE, COMPRESS_16, COMPRESS_16P, COMPRESS_8P = 0, 1, 2, 3

    def __init__(self, nThreads=1, compression=1):
        self.compression = compression

    @staticmethod
    def nshort_pack(nfloats, compression=1):
        return int(lib().klrtm_nshort_pack(nfloats, compression))

    def nshort_volume(self, nz, nx, ny):
        if self.compression == self.COMPRESS_16:
            return nshort_volume(nz, nx, ny)
        return self.nshort_pack(nz, self.compression) * nx * ny

    def compress(self, d_in, d_out, nx, ny, nz):
        c = self.compression
        if c == self.COMPRESS_16:
            check(lib().klrtm_compress16(d_in.ptr, d_out.ptr, nx, ny, nz))
        elif c in (self.COMPRESS_16P, self.COMPRESS_8P):
            check(lib().klrtm_compress_pair(d_in.ptr, d_out.ptr, nx, ny, nz, c))
        else:
            check(lib().klrtm_compress_none(d_in.ptr, d_out.ptr, nx, ny, nz))

    def uncompress(self, d_in, d_out, nx, ny, nz):
        c = self.compression
        if c == self.COMPRESS_16:
            check(lib().klrtm_uncompress16(d_in.ptr, d_out.ptr, nx, ny, nz))
        elif c in (self.COMPRESS_16P, self.COMPRESS_8P):
            check(lib().klrtm_uncompress_pair(d_in.ptr, d_out.ptr, nx, ny, nz, c))
        else:
            check(lib().klrtm_uncompress_none(d_in.ptr, d_out.ptr, nx, ny, nz))


def blockq_nscales(nx, ny, nz):
    return int(lib().klrtm_blockq_nscales(nx, ny, nz))


def blockq_encode(d_in, d_q, d_scales, nx, ny, nz, fmt):
    check(lib().klrtm_blockq_encode(d_in.ptr, d_q.ptr, d_scales.ptr, nx, ny, nz, fmt))


def blockq_decode(d_q, d_scales, d_out, nx, ny, nz, fmt):
    check(lib().klrtm_blockq_decode(d_q.ptr, d_scales.ptr, d_out.ptr, nx, ny, nz, fmt))
