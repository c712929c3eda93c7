"""Oracle RTM pipeline: one shot on the CPU, step by step like the reference.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Restates
  * Receiver::update_zrange + spreadCoeffs   bgprtm/Receiver.cpp:129-151, 306-363
  * Receiver::apply (omp buffer form)        bgprtm/Receiver.cpp:428-475, 1204-1279
  * Receiver::extract_value                  bgprtm/Receiver.cpp:394-426
  * Propagator::modelling time loop          bgprtm/Propagator.cpp:1101-1153
  * Propagator::migration time loop          bgprtm/Propagator.cpp:1210-1255
  * ImagingCondition tapers + crossone       bgprtm/ImagingCondition.cpp:180-226,
                                             bgprtm/ImagingCorrelator.cpp:180-194
  * ImagingCorrelator::run (FFT options)     bgprtm/ImagingCorrelator.cpp:355-439 (oracle/correlator.py)
with the oracle engine for KernelCPU/FdEngine/PML.
"""
import ctypes as C
import math

import numpy as np

from . import correlator
from . import (lib, Engine, spread_coeff, taper_kernel, xy_taper, compress16, uncompress16, apply_symmetry, _pd,
               vel2_q, update2nd_q, fft_size, timeshift)

_pf, _pi = C.POINTER(C.c_float), C.POINTER(C.c_int)


def ricker(nt, dt, fp, dtype=np.float32):
    """Closed-form Ricker, f(t) = (1 - 2 a) exp(-a), a = (pi fp (t - 1.5/fp))^2."""
    t = np.arange(nt, dtype=np.float64) * dt
    tau = t - 1.5 / fp
    a = (math.pi * fp * tau) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(dtype)


def spread_tables(nx, ny, nz, xyz, spread, k1, k2):
    """Per-receiver separable kernels (lx,ly,lz taps and offsets)."""
    xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
    nr = xyz.shape[0]
    izmin, izmax = int(math.floor(xyz[:, 2].min())), int(math.floor(xyz[:, 2].max()))
    spz = spread
    zlimit = min(izmin + 1, nz - 1 - izmax)
    if zlimit < spz:
        spz = max(5, zlimit)
    iz0, iz1 = izmin - spz + 1, izmax + spz
    assert iz0 >= 0 and iz1 < nz, "receiver z range leaves the grid"
    limx = 1 if nx == 1 else max(1, 2 * spread)
    limy = 1 if ny == 1 else max(1, 2 * spread)
    limz = max(1, iz1 - iz0 + 1)
    L = max(limx, limy)
    t = dict(L=L, Lz=limz, lx=np.zeros(nr, np.int32), ly=np.zeros(nr, np.int32), lz=np.zeros(nr, np.int32),
             offx=np.zeros(nr, np.int32), offy=np.zeros(nr, np.int32), offz=np.zeros(nr, np.int32),
             kx=np.zeros((nr, L), np.float32), ky=np.zeros((nr, L), np.float32), kz=np.zeros((nr, limz), np.float32))
    wt = xy_taper(10)
    for r in range(nr):
        ox, kx = spread_coeff(nx, float(xyz[r, 0]), k1, k2, limx)
        oy, ky = spread_coeff(ny, float(xyz[r, 1]), k1, k2, limy)
        oz, kz = spread_coeff(nz, float(xyz[r, 2]), k1, k2, limz)
        kx = taper_kernel(kx, ox, wt, nx) if kx.size else kx
        ky = taper_kernel(ky, oy, wt, ny) if ky.size else ky
        t["offx"][r], t["offy"][r], t["offz"][r] = ox, oy, oz
        t["lx"][r], t["ly"][r], t["lz"][r] = kx.size, ky.size, kz.size
        t["kx"][r, :kx.size], t["ky"][r, :ky.size], t["kz"][r, :kz.size] = kx, ky, kz
    return t


def _tab_args(t):
    return (t["kx"].ctypes.data_as(_pf), t["ky"].ctypes.data_as(_pf), t["kz"].ctypes.data_as(_pf), t["L"], t["L"],
            t["Lz"], t["lx"].ctypes.data_as(_pi), t["ly"].ctypes.data_as(_pi), t["lz"].ctypes.data_as(_pi),
            t["offx"].ctypes.data_as(_pi), t["offy"].ctypes.data_as(_pi), t["offz"].ctypes.data_as(_pi))


def inject(w0, vel, shape, t, vals, iz_mirror=-1):
    """Receiver::apply; with iz_mirror >= 0 the ghost mirror about that plane (bgprtm/Receiver.cpp:457-475)."""
    ny, nx, nz = shape
    vals = np.ascontiguousarray(vals, np.float32)
    lib().orc_inject_ghost(w0.ctypes.data_as(_pf), vel.ctypes.data_as(_pf), nx, ny, nz, vals.size,
                           vals.ctypes.data_as(_pf), *_tab_args(t), int(iz_mirror))


def extract(w, shape, t, iz_mirror=-1):
    """Receiver::extract_value; with iz_mirror >= 0 the ghost mirror (bgprtm/Receiver.cpp:394-426)."""
    ny, nx, nz = shape
    nr = t["lx"].size
    out = np.zeros(nr, np.float32)
    lib().orc_extract_ghost(w.ctypes.data_as(_pf), nx, ny, nz, nr, out.ctypes.data_as(_pf), *_tab_args(t),
                            int(iz_mirror))
    return out


class Shot:
    """One shot: model (record) and rtm (image) in FP32, reference step order."""

    def __init__(self, nx, ny, nz, d, dt, nt, engine, order, model, nbnd, vmax, src_xyz, f_peak, spread, k1, k2,
                 rcv_xyz, snap_every=1, snap_format=0, tmute_frac=0.1, threads=4, free_surface_iz=-1, surface_taper=1,
                 q=None, ghost=None, rho=None, src_array=None):
        self.shape = (ny, nx, nz)
        self.fs, self.fs_taper = free_surface_iz, surface_taper
        # injection mirrors (Propagator::determine_ghost, bgprtm/Propagator.cpp:1538-1541): FREESURFACE
        # mirrors both sides' injection about the surface; ghost = (iz_mirror, source side, receiver
        # side) for SOURCE_GHOST / RECEIVER_GHOST / GHOST, which also mirrors the modelling extraction
        gz, gs, gr = ghost if ghost is not None else (-1, False, False)
        self.src_mirror = free_surface_iz if free_surface_iz >= 0 else (gz if gs else -1)
        self.rcv_mirror = free_surface_iz if free_surface_iz >= 0 else (gz if gr else -1)
        self.ext_mirror = gz if gr else -1
        # Q attenuation: (1/Q volume, Q::coeffs [cq0, cq[order], wq[order]])
        self.q = None if q is None else (np.ascontiguousarray(q[0], np.float32), np.asarray(q[1], np.float32))
        self.rho = None if rho is None else np.ascontiguousarray(rho, np.float32)
        self.nx, self.ny, self.nz, self.dt, self.nt = nx, ny, nz, dt, nt
        self.d = d
        self.model_type, self.snap_every, self.snap_format = model, snap_every, snap_format
        self.eng = Engine(nx, ny, nz, d, d, d, engine, order, np.float32, threads=threads)
        self.nbnd, self.vmax = nbnd, vmax
        dxyz = np.float32(d) * np.float32(d) * np.float32(d)
        self.dt_dxyz = np.float32(1.0) / dxyz
        self.src = spread_tables(nx, ny, nz, np.asarray(src_xyz, np.float32).reshape(1, 3), spread, k1, k2)
        self.rcv = spread_tables(nx, ny, nz, rcv_xyz, spread, k1, k2)
        self.wavelet = ricker(nt, dt, f_peak)
        self.tmute_frac = tmute_frac
        # source array (Propagator::setupReceiver, bgprtm/Propagator.cpp:783-811): positions and delays
        # in seconds; each source injects the wavelet time-shifted by its delay
        self.src_delays = None
        if src_array is not None:
            self.src = spread_tables(nx, ny, nz, np.asarray(src_array[0], np.float32).reshape(-1, 3), spread, k1, k2)
            self.src_delays = np.asarray(src_array[1], np.float64)

    def set_model(self, vel, eps=None, del_=None, pjx=None, pjy=None):
        v = np.asarray(vel, np.float32)
        dt = np.float32(self.dt)
        if self.q is not None:  # Propagator::Vel2, Q branch (bgprtm/Propagator.cpp:725-741)
            self.m, self.invq = v.copy(), self.q[0].copy()
            vel2_q(self.m, self.invq, dt, self.q[1])
        else:
            self.m = ((v * v) * (dt * dt)).astype(np.float32)  # Propagator::Vel2: dt*dt in FP32
        self.eng.set_model(self.m, eps, del_, pjx, pjy)
        self.eng.set_rho(self.rho)
        # FREESURFACE switches the top PML off (bgprtm/FdEngine.cpp:45)
        self.eng.set_boundary(self.nbnd, self.dt, self.vmax, dim_pml=3, top=self.fs < 0)

    def _surface(self, w0):
        """Propagator::apply_symmetry(w0, -1) after the injection (bgprtm/Propagator.cpp:1119-1125,
        1228-1230); the taper length is the z boundary width nzbnd."""
        if self.fs >= 0:
            nzb = self.nbnd[4] if not np.isscalar(self.nbnd) else self.nbnd
            apply_symmetry(w0, self.fs, -1, nzb, self.fs_taper)

    def _src_vals(self, it):
        """Source values of step it (already scaled by dt_dxyz)."""
        if self.src_delays is None:
            return self.wavelet[it:it + 1] * self.dt_dxyz
        if getattr(self, "_src_tr", None) is None or self._src_tr[0] is not self.wavelet:
            nfft = fft_size(self.nt)
            tr = np.array([timeshift(self.wavelet, d / self.dt, nfft) for d in self.src_delays], np.float32)
            self._src_tr = (self.wavelet, tr)
        return self._src_tr[1][:, it] * self.dt_dxyz

    def _fresh(self):
        self.eng.clean_pml()
        z = np.zeros(self.shape, np.float32)
        if self.q is not None:  # Wavefield::cleanMemory (bgprtm/Wavefield.cpp:116-118)
            order = (self.q[1].size - 1) // 2
            self.wq = [np.zeros((order,) + self.shape, np.float32) for _ in range(2)]
        return z.copy(), z.copy(), z.copy()

    def _step(self, w1, w0, wr):
        """Propagator::kernel() + update2nd; with Q the Q leapfrog on the memory-variable set that
        swaps with the wavefields (bgprtm/Propagator.cpp:1789-1816, bgprtm/Wavefield.cpp:130-133)."""
        if self.q is None:
            self.eng.step(self.model_type, w1, w0, wr)
            return
        self.eng.kernel(self.model_type, w1, w0, wr)
        update2nd_q(w0, w1, self.invq, self.wq[0], self.wq[1], self.q[1])
        self.wq.reverse()

    def forward(self, visit):
        """The forward time loop of Propagator::modelling / migration (bgprtm/Propagator.cpp:1101-1153,
        1210-1243): the step, the source injection of trace sample it-1, the free surface; visit(it, w0)
        then sees the new field (where the record is extracted and the snapshot stored)."""
        w0, w1, wr = self._fresh()
        for it in range(self.nt):
            self._step(w1, w0, wr)
            if 0 <= it - 1 < self.nt:
                inject(w0, self.m, self.shape, self.src, self._src_vals(it - 1),
                       self.src_mirror)
            self._surface(w0)
            visit(it, w0)
            w0, w1 = w1, w0
        return w1

    def backward(self, record, visit):
        """The backward loop of Propagator::migration (bgprtm/Propagator.cpp:1199-1255): it runs from
        nt-1 down; the step, the receiver injection of record sample it+1, the free surface; visit(it, w0)
        then sees the new field (where it is imaged)."""
        w0, w1, wr = self._fresh()
        rec = np.ascontiguousarray(record, np.float32)
        for it in range(self.nt - 1, -1, -1):
            self._step(w1, w0, wr)
            ia = it + 1
            if 0 <= ia < self.nt:
                inject(w0, self.m, self.shape, self.rcv, rec[:, ia] * self.dt_dxyz, self.rcv_mirror)
            self._surface(w0)
            visit(it, w0)
            w0, w1 = w1, w0

    def model(self):
        nr = self.rcv["lx"].size
        rec = np.zeros((nr, self.nt), np.float32)

        def visit(it, w0):
            rec[:, it] = extract(w0, self.shape, self.rcv, self.ext_mirror)
        self.last = self.forward(visit)
        return rec

    def _decode(self, snap):
        if self.snap_format == 1:
            return uncompress16(snap, *self.shape)
        # the reference's decoder applies the same edge taper to every codec
        return _edge_taper(snap.reshape(self.shape))

    def rtm(self, record, ic_option=0, kmin=1.0 / 1492.0, dbl=None, high=(0.9, 0.7)):
        """Image and illumination of the record.  ic_option selects ImagingCondition::getImagingOption's
        imaging condition (bgprtm/ImagingCondition.cpp:151-160): 0 the zero-lag product (FP32 sums), 1-3 the
        FFT conditions of oracle.correlator (FP64 sums), which also return the FWI and raw images.  dbl =
        (dz, dx, dy) images on the doubled grid (ImagingCondition::stackImageInterp, :104-141, and the
        post-loop interleave / interpIllum, :415-446): returns image, illumination, FWI image there."""
        if dbl is not None and tuple(dbl) != (1, 1, 1):
            return self._rtm_double(record, ic_option, tuple(dbl), high)
        k = self.snap_every
        ntsav = (self.nt + k - 1) // k
        snaps = [None] * ntsav

        def store(it, w0):
            if it % k == 0 and it // k < ntsav:
                snaps[it // k] = compress16(w0) if self.snap_format == 1 else w0.copy()
        self.forward(store)
        img = np.zeros(self.shape, np.float32)
        ill = np.zeros(self.shape, np.float32)
        # ImagingCondition::calTapes rounds an FP32 quotient (bgprtm/ImagingCondition.cpp:185)
        f32 = np.float32
        ntaperT = int(np.rint(f32(self.tmute_frac) * (f32(self.nt) * f32(self.dt)) / (f32(k) * f32(self.dt))))
        wtT = [np.float32(0.5) * (np.float32(1) - np.cos(np.float32(np.pi) * np.float32(i) / np.float32(ntaperT)))
               for i in range(ntaperT)]

        def image(it, w0):
            if it % k == 0 and it // k < ntsav:
                s = it // k
                S = self._decode(snaps[s])
                R = uncompress16(compress16(w0), *self.shape) if self.snap_format == 1 else _edge_taper(w0)
                if s < ntaperT:
                    S = (S * wtT[s]).astype(np.float32)
                ill[...] += S * S
                if ic_option == 0:
                    img[...] += S * R
                else:
                    acc[0], acc[1], acc[2] = correlator.run(ic_option, S, R, (self.d, self.d, self.d), kmin, *acc)
        acc = [np.zeros(self.shape), np.zeros(self.shape), np.zeros(self.shape)]
        self.backward(record, image)
        self.snaps = snaps  # the stored snapshots (COMPRESS_16 codes or FP32 volumes), for cache checks
        if ic_option == 0:
            return img, ill
        if ic_option == 2:   # the post-loop cos and sin terms of INVERSION2 (bgprtm/ImagingCondition.cpp:427-430)
            acc[0] = correlator.costerm(acc[0], (self.d, self.d, self.d))
            acc[1] = correlator.sinterm(acc[1], acc[0])
        return acc[0], ill, acc[1], acc[2]


    def _rtm_double(self, record, ic_option, dbl, high):
        k = self.snap_every
        ntsav = (self.nt + k - 1) // k
        snaps = [None] * ntsav

        def store(it, w0):
            if it % k == 0 and it // k < ntsav:
                snaps[it // k] = compress16(w0) if self.snap_format == 1 else w0.copy()
        self.forward(store)
        f32 = np.float32
        ntaperT = int(np.rint(f32(self.tmute_frac) * (f32(self.nt) * f32(self.dt)) / (f32(k) * f32(self.dt))))
        wtT = [np.float32(0.5) * (np.float32(1) - np.cos(np.float32(np.pi) * np.float32(i) / np.float32(ntaperT)))
               for i in range(ntaperT)]
        nloop = len(correlator.double_copies(*dbl))
        acc = [np.zeros((nloop,) + self.shape), np.zeros((nloop,) + self.shape)]
        ill = np.zeros(self.shape, np.float32)

        def image(it, w0):
            if it % k == 0 and it // k < ntsav:
                s = it // k
                S = self._decode(snaps[s])
                R = uncompress16(compress16(w0), *self.shape) if self.snap_format == 1 else _edge_taper(w0)
                if s < ntaperT:
                    S = (S * wtT[s]).astype(np.float32)
                Sh, acc[0], acc[1] = correlator.run_double(ic_option, S, R, (self.d, self.d, self.d), dbl, high, *acc)
                Sh = Sh.astype(np.float32)
                ill[...] += Sh * Sh
        self.backward(record, image)
        fwi = correlator.interleave(acc[1], dbl) if ic_option == 1 else None
        return correlator.interleave(acc[0], dbl), correlator.interp_illum(ill, dbl), fwi


class Shot64:
    """The same shot in FP64 throughout (engine, injection, extraction, snapshots, imaging): the
    reference algorithm without rounding, against which reduced-precision images are measured.
    The spread taps keep their FP32 values (they are part of the problem definition)."""

    def __init__(self, nx, ny, nz, d, dt, nt, engine, order, model, nbnd, vmax, src_xyz, f_peak, spread, k1, k2,
                 rcv_xyz, snap_every=1, tmute_frac=0.1, threads=4):
        self.shape = (ny, nx, nz)
        self.nx, self.ny, self.nz, self.dt, self.nt = nx, ny, nz, dt, nt
        self.model_type, self.snap_every, self.tmute_frac = model, snap_every, tmute_frac
        self.eng = Engine(nx, ny, nz, d, d, d, engine, order, np.float64, threads=threads)
        self.nbnd, self.vmax = nbnd, vmax
        self.dt_dxyz = 1.0 / (float(d) ** 3)
        self.src = spread_tables(nx, ny, nz, np.asarray(src_xyz, np.float32).reshape(1, 3), spread, k1, k2)
        self.rcv = spread_tables(nx, ny, nz, rcv_xyz, spread, k1, k2)
        self.wavelet = ricker(nt, dt, f_peak, dtype=np.float64)

    def set_model(self, vel, eps=None, del_=None, pjx=None, pjy=None):
        v = np.asarray(vel, np.float64)
        self.m = v * v * (self.dt * self.dt)
        f64 = [None if a is None else np.asarray(a, np.float64) for a in (eps, del_, pjx, pjy)]
        self.eng.set_model(self.m, *f64)
        self.eng.set_boundary(self.nbnd, self.dt, self.vmax, dim_pml=3)

    def _inject(self, w0, t, vals):
        ny, nx, nz = self.shape
        vals = np.ascontiguousarray(vals, np.float64)
        lib().orc_inject_d(w0.ctypes.data_as(_pd), self.m.ctypes.data_as(_pd), nx, ny, nz, vals.size,
                           vals.ctypes.data_as(_pd), *_tab_args(t))

    def _extract(self, w, t):
        ny, nx, nz = self.shape
        out = np.zeros(t["lx"].size, np.float64)
        lib().orc_extract_d(w.ctypes.data_as(_pd), nx, ny, nz, out.size, out.ctypes.data_as(_pd), *_tab_args(t))
        return out

    def _fresh(self):
        self.eng.clean_pml()
        z = np.zeros(self.shape, np.float64)
        return z.copy(), z.copy(), z.copy()

    def forward(self, visit):
        """Shot.forward in FP64."""
        w0, w1, wr = self._fresh()
        for it in range(self.nt):
            self.eng.step(self.model_type, w1, w0, wr)
            if 0 <= it - 1 < self.nt:
                self._inject(w0, self.src, self.wavelet[it - 1:it] * self.dt_dxyz)
            visit(it, w0)
            w0, w1 = w1, w0
        return w1

    def backward(self, record, visit):
        """Shot.backward in FP64."""
        w0, w1, wr = self._fresh()
        rec = np.asarray(record, np.float64)
        for it in range(self.nt - 1, -1, -1):
            self.eng.step(self.model_type, w1, w0, wr)
            ia = it + 1
            if 0 <= ia < self.nt:
                self._inject(w0, self.rcv, rec[:, ia] * self.dt_dxyz)
            visit(it, w0)
            w0, w1 = w1, w0

    def model(self):
        rec = np.zeros((self.rcv["lx"].size, self.nt), np.float64)

        def visit(it, w0):
            rec[:, it] = self._extract(w0, self.rcv)
        self.forward(visit)
        return rec

    def rtm(self, record, ic_option=0, kmin=1.0 / 1492.0):
        """Image and illumination of the record.  ic_option selects ImagingCondition::getImagingOption's
        imaging condition (bgprtm/ImagingCondition.cpp:151-160): 0 the zero-lag product (FP32 sums), 1-3 the
        FFT conditions of oracle.correlator (FP64 sums), which also return the FWI and raw images."""
        k = self.snap_every
        ntsav = (self.nt + k - 1) // k
        snaps = [None] * ntsav

        def store(it, w0):
            if it % k == 0 and it // k < ntsav:
                snaps[it // k] = w0.copy()
        self.forward(store)
        img = np.zeros(self.shape, np.float64)
        ill = np.zeros(self.shape, np.float64)
        f32 = np.float32  # the taper length rounds an FP32 quotient, as in Shot (a count, not a precision)
        ntaperT = int(np.rint(f32(self.tmute_frac) * (f32(self.nt) * f32(self.dt)) / (f32(k) * f32(self.dt))))

        def image(it, w0):
            if it % k == 0 and it // k < ntsav:
                s = it // k
                S = _edge_taper(snaps[s], np.float64)
                R = _edge_taper(w0, np.float64)
                if s < ntaperT:
                    S = S * (0.5 * (1 - math.cos(math.pi * s / ntaperT)))
                img[...] += S * R
                ill[...] += S * S
        self.backward(record, image)
        return img, ill


def _edge_taper(v, dtype=np.float32):
    """Edge taper of WaveFieldCompress::uncompress (20 x/y cells, 10 z cells), in dtype."""
    ny, nx, nz = v.shape
    out = np.array(v, dtype, copy=True)
    if dtype == np.float32:
        wt = (0.5 * (1 - np.cos(np.float32(np.pi) * np.arange(20, dtype=np.float32) / np.float32(20)))).astype(np.float32)
        wz = (0.5 * (1 - np.cos(np.float32(np.pi) * np.arange(10, dtype=np.float32) / np.float32(10)))).astype(np.float32)
    else:
        wt = 0.5 * (1 - np.cos(np.pi * np.arange(20) / 20.0))
        wz = 0.5 * (1 - np.cos(np.pi * np.arange(10) / 10.0))
    if ny > 40:
        out[:20] *= wt[:, None, None]
        out[ny - 20:] *= wt[::-1, None, None]
    if nx > 40:
        out[:, :20] *= wt[None, :, None]
        out[:, nx - 20:] *= wt[None, ::-1, None]
    if nz > 20:
        out[:, :, :10] *= wz[None, None, :]
        out[:, :, nz - 10:] *= wz[None, None, ::-1]
    return out
