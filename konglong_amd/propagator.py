"""Propagator: one RTM shot on the GPU (bgprtm/Propagator.h:100-104 rtm / mod)."""
import ctypes as C

import numpy as np

from ._lib import lib, check, require, fptr, dptr, ShotConfig, ENGINE_FD, ISO, SNAP_FP32


def shot_config(nx, ny, nz, d, dt, nt, engine=ENGINE_FD, order=8, model=ISO, nbnd=20, vmax=None, src_xyz=None,
                f_peak=15.0, spread=10, khpass=0.3, khcut=0.3, nrcv=0, snap_every=1, snap_format=SNAP_FP32,
                rec_through_codec=0, tmute_frac=0.1, ntaper_z=0, fuse_step=1, pml_eps=1e-3, pml3d=1,
                pml3d_scaler=0.3, top_pml=1, snap_slots=0, ic_fp32_only=0, free_surface_iz=-1, surface_taper=1):
    c = ShotConfig()
    c.nx, c.ny, c.nz = nx, ny, nz
    dx, dy, dz = (d, d, d) if np.isscalar(d) else d
    c.dx, c.dy, c.dz, c.dt, c.nt = dx, dy, dz, dt, nt
    c.engine, c.order, c.model = engine, order, model
    nb = [nbnd] * 6 if np.isscalar(nbnd) else list(nbnd)
    for i in range(6):
        c.nbnd[i] = nb[i]
    c.vmax = vmax if vmax is not None else 4500.0
    c.pml_eps, c.pml3d_scaler, c.pml3d, c.top_pml = pml_eps, pml3d_scaler, pml3d, top_pml
    sx = src_xyz if src_xyz is not None else ((nx - 1) / 2, (ny - 1) / 2, (nz - 1) / 2)
    for i in range(3):
        c.src_xyz[i] = sx[i]
    c.f_peak = f_peak
    c.src_spread = c.rcv_spread = spread
    c.khpass, c.khcut = khpass, khcut
    c.nrcv = nrcv
    c.snap_every, c.snap_format, c.rec_through_codec = snap_every, snap_format, rec_through_codec
    c.tmute_frac, c.ntaper_z, c.fuse_step = tmute_frac, ntaper_z, fuse_step
    c.snap_slots = snap_slots
    c.ic_fp32_only = ic_fp32_only
    c.free_surface_iz, c.surface_taper = free_surface_iz, surface_taper
    return c


class Propagator:
    def __init__(self, cfg):
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().klrtm_propagator_create(C.byref(h), C.byref(cfg)))
        self.h = h.value
        self.shape = (cfg.ny, cfg.nx, cfg.nz)

    def set_model(self, vel, eps=None, del_=None, pjx=None, pjy=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (vel, eps, del_, pjx, pjy)]
        check(lib().klrtm_propagator_set_model(self.h, *[fptr(a) for a in arrs]))

    def set_rho(self, rho):
        """Density (Model::useRho): the gradient is divided by rho and each divergence term scaled by
        vel*rho; ISO takes the first-derivative form.  None switches it off."""
        r = None if rho is None else np.ascontiguousarray(rho, np.float32)
        require(r is None or r.shape == self.shape, "density volume of the propagator's shape")
        check(lib().klrtm_propagator_set_rho(self.h, fptr(r)))

    def set_q(self, invq, coeffs):
        """Q attenuation (Model::useQ): the 1/Q volume and Q::coeffs [cq0, cq[order], wq[order]]; call
        before set_model.  invq None switches Q off."""
        if invq is None:
            check(lib().klrtm_propagator_set_q(self.h, None, 0, None))
            return
        q = np.ascontiguousarray(invq, np.float32)
        cf = np.ascontiguousarray(coeffs, np.float32)
        require(q.shape == self.shape and cf.size % 2 == 1, "1/Q volume of the propagator's shape, 2 order + 1 coefficients")
        check(lib().klrtm_propagator_set_q(self.h, fptr(q), (cf.size - 1) // 2, fptr(cf)))

    def set_ghost(self, iz_mirror, source_side=True, receiver_side=True):
        """SOURCE_GHOST / RECEIVER_GHOST / GHOST surfaces: mirror the chosen sides' injection (and with
        receiver_side the modelling extraction) about iz_mirror."""
        check(lib().klrtm_propagator_set_ghost(self.h, int(iz_mirror), int(source_side), int(receiver_side)))

    def set_snapshot_cache(self, path=None, mem_limit_gb=0.0, keep=False):
        """CacheFile: stream the snapshots of rtm to host memory (when they fit mem_limit_gb) or to the
        file at path instead of holding them in HBM."""
        check(lib().klrtm_propagator_set_snapshot_cache(self.h, None if path is None else str(path).encode(),
                                                        float(mem_limit_gb), int(keep)))

    def snapshot_cache_info(self):
        m, st, ns = C.c_int(), C.c_size_t(), C.c_int()
        check(lib().klrtm_propagator_snapshot_cache_info(self.h, C.byref(m), C.byref(st), C.byref(ns)))
        return dict(in_memory=bool(m.value), stride=st.value, nsnap=ns.value)

    def set_source_array(self, xyz, delays_s):
        """Source array (Propagator::setupReceiver): the wavelet injected at each position, delayed by
        its delay in seconds (FFT time shift)."""
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        dl = np.ascontiguousarray(delays_s, np.float32).reshape(-1)
        require(dl.size == xyz.shape[0], "one delay per source")
        check(lib().klrtm_propagator_set_source_array(self.h, xyz.shape[0], fptr(xyz), fptr(dl)))

    def set_source_wavelet(self, wavelet):
        w = np.ascontiguousarray(wavelet, np.float32)
        require(w.shape == (self.cfg.nt,), f"wavelet of {self.cfg.nt} samples")
        check(lib().klrtm_propagator_set_source_wavelet(self.h, fptr(w)))

    def set_receivers(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        require(xyz.shape[0] == self.cfg.nrcv, f"{self.cfg.nrcv} receiver positions")
        check(lib().klrtm_propagator_set_receivers(self.h, fptr(xyz)))

    def model(self, want_record=True):
        rec = np.zeros((self.cfg.nrcv, self.cfg.nt), np.float32) if want_record else None
        check(lib().klrtm_propagator_model(self.h, fptr(rec)))
        return rec

    def set_imaging(self, option, kmin=1.0 / 1492.0):
        """The imaging condition (LowcutFilter): IC_IMAGE1, IC_INVERSION, IC_INVERSION2 or IC_ALLIN3; kmin is
        minFreq / vMin (bgprtm/ImagingCondition.cpp:393-396)."""
        check(lib().klrtm_propagator_set_imaging(self.h, int(option), float(kmin)))
        self.ic_option = int(option)

    def set_zgrid(self, zgrid=None):
        """IRREGULAR grid: the depth of every z index (both engines); None returns to the regular 1/dz."""
        z = None if zgrid is None else np.ascontiguousarray(zgrid, np.float32)
        if z is not None and z.shape != (self.shape[2],):
            raise ValueError(f"zgrid needs {self.shape[2]} depths, got shape {z.shape}")
        check(lib().klrtm_propagator_set_zgrid(self.h, None if z is None else z.ctypes.data))

    def set_ic_double(self, dz=2, dx=2, dy=2, high_z=0.9, high_xy=0.7):
        """Image on a grid doubled along z, x, y (ic_double_z/x/y with highTapeZ / highTapeXY): 3D grids
        double all three axes, 2D grids z or z and x; (1, 1, 1) restores the propagation grid.  rtm() then
        returns the image, FWI image and illumination on the doubled grid (FP32)."""
        if self.shape[0] == 1:
            dy = 1
        check(lib().klrtm_propagator_set_ic_double(self.h, int(dz), int(dx), int(dy), float(high_z), float(high_xy)))
        self.ic_dbl = (int(dz), int(dx), int(dy))

    def double_images(self):
        """Device pointers of the doubled-grid image, FWI image and illumination, their shape and nloop."""
        img, fwi, ill, nl = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_int()
        check(lib().klrtm_propagator_ic_double_images(self.h, C.byref(img), C.byref(fwi), C.byref(ill), C.byref(nl)))
        dz, dx, dy = getattr(self, "ic_dbl", (1, 1, 1))
        ny, nx, nz = self.shape
        return dict(image=img.value, fwi=fwi.value, illum=ill.value, shape=(ny * dy, nx * dx, nz * dz), nloop=nl.value)

    def rtm(self, record=None, want=("image", "illum", "image64", "illum64")):
        n = self.shape
        r = None if record is None else np.ascontiguousarray(record, np.float32)
        if getattr(self, "ic_dbl", (1, 1, 1)) != (1, 1, 1):
            check(lib().klrtm_propagator_rtm(self.h, fptr(r), None, None, None, None))
            d = self.double_images()
            out = {}
            for name in ("image", "fwi", "illum"):
                if d[name]:
                    a = np.empty(d["shape"], np.float32)
                    check(lib().klrtm_memcpy_d2h(a.ctypes.data, d[name], a.nbytes))
                    out[name] = a
            return out
        img = np.zeros(n, np.float32) if "image" in want else None
        ill = np.zeros(n, np.float32) if "illum" in want else None
        img64 = np.zeros(n, np.float64) if "image64" in want else None
        ill64 = np.zeros(n, np.float64) if "illum64" in want else None
        check(lib().klrtm_propagator_rtm(self.h, fptr(r), fptr(img), fptr(ill), dptr(img64), dptr(ill64)))
        out = dict(image=img, illum=ill, image64=img64, illum64=ill64)
        if getattr(self, "ic_option", 0):   # the FWI and raw images of the FFT imaging conditions
            fwi, raw = C.c_void_p(), C.c_void_p()
            check(lib().klrtm_propagator_ic_images(self.h, C.byref(fwi), C.byref(raw)))
            for name, ptr in (("fwi", fwi), ("raw", raw)):
                a = np.empty(n, np.float32)
                check(lib().klrtm_memcpy_d2h(a.ctypes.data, ptr.value, a.nbytes))
                out[name] = a
        return out

    def timings(self):
        f, b, i = C.c_double(), C.c_double(), C.c_double()
        check(lib().klrtm_propagator_timings(self.h, C.byref(f), C.byref(b), C.byref(i)))
        return dict(forward_ms=f.value, backward_ms=b.value, ic_ms=i.value)

    def forward_steps(self, it0, nsteps, store_snapshots=False):
        check(lib().klrtm_propagator_forward_steps(self.h, it0, nsteps, int(store_snapshots)))

    def backward_steps(self, it0, nsteps, image=True):
        check(lib().klrtm_propagator_backward_steps(self.h, it0, nsteps, int(image)))

    def set_graph(self, on=True):
        """Run whole snapshot-interval groups of forward_steps / backward_steps as replays of captured
        hipGraphs (a device step counter supplies the time index); results are unchanged."""
        check(lib().klrtm_propagator_set_graph(self.h, int(on)))

    def graph_stats(self):
        g, r = C.c_int(), C.c_long()
        check(lib().klrtm_propagator_graph_stats(self.h, C.byref(g), C.byref(r)))
        return g.value, r.value

    def set_record(self, record):
        r = np.ascontiguousarray(record, np.float32)
        require(r.shape == (self.cfg.nrcv, self.cfg.nt), f"record of shape ({self.cfg.nrcv}, {self.cfg.nt})")
        check(lib().klrtm_propagator_set_record(self.h, fptr(r)))

    def reset(self):
        check(lib().klrtm_propagator_reset(self.h))

    def wavefield(self):
        """Current w1 (the latest time level after a step's swap), downloaded."""
        w0, w1 = C.c_void_p(), C.c_void_p()
        check(lib().klrtm_propagator_wavefields(self.h, C.byref(w0), C.byref(w1)))
        out = np.empty(self.shape, np.float32)
        check(lib().klrtm_memcpy_d2h(out.ctypes.data, w1.value, out.nbytes))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().klrtm_propagator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
