"""Receiver / ExtractReceiverData on the GPU (bgprtm/Receiver.h, bgprtm/ExtractReceiverData.h)."""
import ctypes as C

import numpy as np

from ._lib import lib, check, require, fptr, iptr


def spread_coeff(n, x, k1, k2, nlimit):
    """Util::CompactOrmsbySpreadCoeff: returns (offset, taps)."""
    coeff = np.zeros(max(nlimit, 1), np.float32)
    nl, off = C.c_int(nlimit), C.c_int(0)
    check(lib().klrtm_spread_coeff(n, fptr(coeff), x, k1, k2, C.byref(nl), C.byref(off)))
    return off.value, coeff[: nl.value].copy()


class Receiver:
    """nr traces at grid coordinates xyz (nr x 3, index units) with spread kernels."""

    def __init__(self, nx, ny, nz, xyz, spread, k1, k2, nt, dt_dxyz):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        self.nr, self.nt = xyz.shape[0], nt
        h = C.c_void_p()
        check(lib().klrtm_receiver_create(C.byref(h), nx, ny, nz, self.nr, fptr(xyz), spread, k1, k2, nt, dt_dxyz))
        self.h = h.value

    def set_traces(self, traces):
        t = np.ascontiguousarray(traces, np.float32)
        require(t.shape == (self.nr, self.nt), f"traces of shape ({self.nr}, {self.nt})")
        check(lib().klrtm_receiver_set_traces(self.h, fptr(t)))

    def apply(self, w0, it, vel, scaler0=1.0, iz_mirror=-1):
        """Receiver::apply: w0 += trace[:, it]*scaler0*dt_dxyz*kx*ky*kz*vel; with iz_mirror >= 0 the
        ghost mirror about that plane is subtracted too."""
        if iz_mirror >= 0:
            check(lib().klrtm_receiver_apply_ghost(self.h, w0.ptr, it, vel.ptr, scaler0, iz_mirror))
        else:
            check(lib().klrtm_receiver_apply(self.h, w0.ptr, it, vel.ptr, scaler0))

    def extract(self, w, jt, iz_mirror=-1):
        if iz_mirror >= 0:
            check(lib().klrtm_receiver_extract_ghost(self.h, w.ptr, jt, iz_mirror))
        else:
            check(lib().klrtm_receiver_extract(self.h, w.ptr, jt))

    def fk_spread(self, dx, dy, dt, vsurf, trueamp=1, z_depth=0.0, deghost=0, sign=-1, ntfft=None, flcut=0.0,
                  flpass=0.0, zerofreq_threshold=0.1):
        """Receiver::spread_sdomain: FK true amplitude / deghost of the traces set last; apply() then injects
        the surface wavefield it leaves (sign -1 receiver side, +1 source side)."""
        if ntfft is None:
            ntfft = int(lib().klrtm_fft_size(int(round(self.nt * 1.5))))
        check(lib().klrtm_receiver_fk_spread(self.h, dx, dy, dt, ntfft, vsurf, int(trueamp), z_depth, int(deghost),
                                             int(sign), flcut, flpass, zerofreq_threshold))
        return ntfft

    def fk_data(self, shape):
        """The FK surface wavefield [nt][ny][nx] (shape = (ny, nx)), downloaded."""
        p = C.c_void_p()
        check(lib().klrtm_receiver_fk_data(self.h, C.byref(p)))
        out = np.empty((self.nt,) + tuple(shape), np.float32)
        check(lib().klrtm_memcpy_d2h(out.ctypes.data, p.value, out.nbytes))
        return out

    def record(self):
        out = np.zeros((self.nr, self.nt), np.float32)
        check(lib().klrtm_receiver_get_record(self.h, fptr(out)))
        return out

    def tables(self):
        n = self.nr
        maxlen = np.zeros(2, np.int32)
        check(lib().klrtm_receiver_get_tables(self.h, None, None, None, None, None, None, None, None, None, iptr(maxlen)))
        L, Lz = int(maxlen[0]), int(maxlen[1])
        ints = [np.zeros(n, np.int32) for _ in range(6)]
        kx, ky, kz = np.zeros(n * L, np.float32), np.zeros(n * L, np.float32), np.zeros(n * Lz, np.float32)
        check(lib().klrtm_receiver_get_tables(self.h, *[iptr(a) for a in ints], fptr(kx), fptr(ky), fptr(kz),
                                              iptr(maxlen)))
        lx, ly, lz, offx, offy, offz = ints
        return dict(lx=lx, ly=ly, lz=lz, offx=offx, offy=offy, offz=offz, kx=kx.reshape(n, L), ky=ky.reshape(n, L),
                    kz=kz.reshape(n, Lz), L=L, Lz=Lz)

    def close(self):
        if getattr(self, "h", None):
            lib().klrtm_receiver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
