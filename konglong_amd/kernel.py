"""KernelCPU: the bgprtm per-step physics API (bgprtm/KernelCPU.h:19-35) on the GPU.

Wavefield and model volumes are ``DeviceArray`` objects (device memory);
the arithmetic runs in the HIP kernels of libkonglong_rtm.so.
"""
import ctypes as C

import numpy as np

from ._lib import lib, check, ENGINE_FD, ENGINE_CFD, ISO, VTI, TTI  # noqa: F401


def _p(d):
    return None if d is None else d.ptr


class KernelCPU:
    def __init__(self, nx, ny, nz, dx, dy, dz, engine=ENGINE_FD, order=8):
        h = C.c_void_p()
        check(lib().klrtm_kernel_create(C.byref(h), nx, ny, nz, dx, dy, dz, engine, order))
        self.h = h.value
        self.nx, self.ny, self.nz = nx, ny, nz
        self._keep = ()

    def setModel(self, vel, rho=None, del_=None, eps=None, pjx=None, pjy=None):
        self._keep = (vel, rho, del_, eps, pjx, pjy)
        check(lib().klrtm_kernel_set_model(self.h, _p(vel), _p(rho), _p(del_), _p(eps), _p(pjx), _p(pjy)))

    def setPmlOptions(self, pml_eps=1e-3, pml3d=True, pml3d_scaler=0.3, top_pml_flag=True):
        check(lib().klrtm_kernel_set_pml_options(self.h, pml_eps, int(pml3d), pml3d_scaler, int(top_pml_flag)))

    def setBoundary(self, nxbnd1, nxbnd2, nybnd1, nybnd2, nzbnd1, nzbnd2, dt, vmax):
        check(lib().klrtm_kernel_set_boundary(self.h, nxbnd1, nxbnd2, nybnd1, nybnd2, nzbnd1, nzbnd2, dt, vmax))

    def cleanPMLMemory(self):
        check(lib().klrtm_kernel_clean_pml(self.h))

    def ISO(self, p0, pb, pr):
        check(lib().klrtm_kernel_iso(self.h, p0.ptr, pb.ptr, pr.ptr))

    def VTI(self, p0, pb, pr):
        check(lib().klrtm_kernel_vti(self.h, p0.ptr, pb.ptr, _p(pr)))

    def TTI(self, p0, pb, pr):
        check(lib().klrtm_kernel_tti(self.h, p0.ptr, pb.ptr, _p(pr)))

    def step(self, model, w1, w0, wr):
        """kernel() + update2nd() fused: w0 <- 2 w1 - (w0 - vel*Op(w1))."""
        check(lib().klrtm_kernel_step(self.h, model, w1.ptr, w0.ptr, _p(wr)))

    def set_tile(self, tz=0, tx=0, ychunk=0):
        check(lib().klrtm_kernel_set_tile(self.h, tz, tx, ychunk))

    def set_zgrid(self, zgrid=None):
        """IRREGULAR grid: the depth of every z index (FdEngine_fd::setJacob); None returns to 1/dz."""
        if zgrid is None:
            check(lib().klrtm_kernel_set_zgrid(self.h, None))
            return
        z = np.ascontiguousarray(zgrid, np.float32)
        if z.shape != (self.nz,):
            raise ValueError(f"zgrid needs {self.nz} depths, got shape {z.shape}")
        check(lib().klrtm_kernel_set_zgrid(self.h, z.ctypes.data))

    def set_overlap(self, on=True):
        """Compact engine: independent line passes of a step on two streams (bit-identical results)."""
        check(lib().klrtm_kernel_set_overlap(self.h, int(on)))

    def get_tile(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().klrtm_kernel_get_tile(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def close(self):
        if getattr(self, "h", None):
            lib().klrtm_kernel_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FdEngine:
    """FdEngine volume operations (bgprtm/FdEngine.h:45-119) on device arrays.

    ``laplacian(in, out)`` (FD engines) and ``deriv1(axis, in, out, iacc, isign)``
    (axis 0 z, 1 x, 2 y; iacc 1 accumulates ``out -= vel*D(in)`` with the model set
    through ``setModel``).  The PML state is the engine's own, as in the reference.
    """

    ENGINE = ENGINE_FD

    def __init__(self, nx, ny, nz, dx, dy, dz, order=8):
        self._k = KernelCPU(nx, ny, nz, dx, dy, dz, self.ENGINE, order)

    def setModel(self, vel, rho=None, del_=None, eps=None, pjx=None, pjy=None):
        self._k.setModel(vel, rho, del_, eps, pjx, pjy)

    def setPmlOptions(self, **kw):
        self._k.setPmlOptions(**kw)

    def setBoundary(self, *a):
        self._k.setBoundary(*a)

    def cleanPMLMemory(self):
        self._k.cleanPMLMemory()

    def laplacian(self, u, out):
        check(lib().klrtm_kernel_laplacian(self._k.h, u.ptr, out.ptr))

    def deriv1(self, axis, u, out, iacc=0, isign=1):
        check(lib().klrtm_kernel_deriv1(self._k.h, axis, u.ptr, out.ptr, iacc, isign))

    def dz1(self, u, out, iacc=0, isign=1):
        self.deriv1(0, u, out, iacc, isign)

    def dx1(self, u, out, iacc=0, isign=1):
        self.deriv1(1, u, out, iacc, isign)

    def dy1(self, u, out, iacc=0, isign=1):
        self.deriv1(2, u, out, iacc, isign)

    def close(self):
        self._k.close()


class FdEngine_fd(FdEngine):
    ENGINE = ENGINE_FD


class FdEngine_cfd(FdEngine):
    ENGINE = ENGINE_CFD

    def __init__(self, nx, ny, nz, dx, dy, dz, order=10):
        super().__init__(nx, ny, nz, dx, dy, dz, 10)

