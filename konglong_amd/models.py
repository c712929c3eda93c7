"""Seeded synthetic earth models and the CFL time step.

The reference loads models from files (ModelLoader) and derives dt from the
CFL rule of ModelPrepare::calcDeltaTime (bgprtm/ModelPrepare.cpp:647-870);
these generators build the benchmark workloads' models directly.
"""
import math

import numpy as np

from ._lib import ENGINE_FD, ISO, VTI, TTI


def gaussian_smooth_noise(shape, sigma, seed, dtype=np.float32):
    """White noise filtered by a 3D Gaussian of std `sigma` cells (spectral), zero mean, unit std."""
    import scipy.fft as sfft
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape, dtype=np.float32)
    F = sfft.rfftn(f, workers=-1)
    for ax, n in enumerate(shape):
        k = sfft.rfftfreq(n) if ax == len(shape) - 1 else sfft.fftfreq(n)
        g = np.exp(-0.5 * (2 * np.pi * k * sigma) ** 2).astype(np.float32)
        sh = [1] * len(shape)
        sh[ax] = g.size
        F *= g.reshape(sh)
    out = sfft.irfftn(F, s=shape, workers=-1).astype(dtype)
    out -= out.mean()
    out /= out.std() + 1e-30
    return out


def _to_range(f, lo, hi):
    a, b = float(f.min()), float(f.max())
    return (lo + (hi - lo) * (f - a) / (b - a + 1e-30)).astype(np.float32)


def layered_model(n=128, d=10.0, seed=101, nlayers=6, vlo=1500.0, vhi=4500.0, anomaly_sigma=8.0, vclip=4400.0):
    """Flat layers with random depths and velocities plus one +-10% Gaussian anomaly.

    Velocities above `vclip` are clipped so FD8 at dt = 1 ms stays inside its
    3D CFL limit (v dt/dx <= 0.4529 for dx = 10 m).  Array layout (ny, nx, nz).
    """
    rng = np.random.default_rng(seed)
    depths = np.sort(rng.uniform(0.0, n * d, nlayers - 1))
    vels = rng.uniform(vlo, vhi, nlayers)
    z = np.arange(n) * d
    prof = vels[np.searchsorted(depths, z)]
    v = np.broadcast_to(prof[None, None, :], (n, n, n)).astype(np.float32).copy()
    c = rng.uniform(n / 4, 3 * n / 4, 3)
    amp = 0.1 * (1 if rng.random() < 0.5 else -1)
    y, x, zz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    g = np.exp(-((y - c[0]) ** 2 + (x - c[1]) ** 2 + (zz - c[2]) ** 2) / (2 * anomaly_sigma ** 2))
    v *= (1 + amp * g).astype(np.float32)
    return np.minimum(v, vclip).astype(np.float32)


def smooth_vti_model(shape=(400, 400, 400), sigma=20.0, seed=202):
    """Vp: Gaussian-filtered noise mapped to 1500-5000 m/s; eps 0-0.3, delta 0-0.15 (same smoothing)."""
    vp = _to_range(gaussian_smooth_noise(shape, sigma, seed), 1500.0, 5000.0)
    eps = _to_range(gaussian_smooth_noise(shape, sigma, seed + 1), 0.0, 0.3)
    dl = _to_range(gaussian_smooth_noise(shape, sigma, seed + 2), 0.0, 0.15)
    return vp, eps, dl


def tilt_axes(shape, sigma, seed):
    """pjx, pjy from smooth tilt in 0-45 deg and azimuth in 0-360 deg (bgprtm/ModelLoader.cpp:303-321)."""
    th = _to_range(gaussian_smooth_noise(shape, sigma, seed), 0.0, 45.0) * np.float32(np.pi / 180)
    ph = _to_range(gaussian_smooth_noise(shape, sigma, seed + 1), 0.0, 360.0) * np.float32(np.pi / 180)
    st = np.sin(th)
    return (-st * np.cos(ph)).astype(np.float32), (-st * np.sin(ph)).astype(np.float32)


def fd_k1max(order):
    """sqrt of the Nyquist symbol of the Taylor 2nd-derivative stencil, x1.01 (bgprtm/FdEngine_fd.cpp:69-81)."""
    M = order // 2
    c = []
    s = 0.0
    for j in range(1, M + 1):
        r = 1.0
        for k in range(1, j + 1):
            r *= (M - j + k) / (M + k)
        a = 2.0 * (1 if j % 2 else -1) * r / (j * j)
        c.append(a)
        s += a
    c0 = -2.0 * s
    sym = -c0 + sum(2 * cj * (-1) ** (j + 1) for j, cj in enumerate(c, start=1))
    return math.sqrt(sym) * 1.01


def cfl_dt(vel, d, engine=ENGINE_FD, order=8, model=ISO, eps=None, dl=None, pjx=None, pjy=None):
    """dt from the CFL rule of ModelPrepare::calcDeltaTime without density (bgprtm/ModelPrepare.cpp:701-718 ISO,
    :747-765 VTI, :826-870 TTI).  TTI needs the tilt axes pjx, pjy: the grid wavenumber vector is rotated into
    the symmetry frame and the bound is the largest qP rate over the 7 non-empty axis subsets."""
    dx, dy, dz = (d, d, d) if np.isscalar(d) else d
    k1max = fd_k1max(order) if engine == ENGINE_FD else math.pi
    rdx2, rdy2, rdz2 = 1.0 / dx ** 2, 1.0 / dy ** 2, 1.0 / dz ** 2
    v = np.asarray(vel, np.float64)
    if model == ISO:
        lam = v * math.sqrt(rdx2 + rdy2 + rdz2)
        return float(np.min(2.0 / k1max / lam))
    if model == TTI:
        if pjx is None or pjy is None:
            raise ValueError("the TTI CFL rule needs the tilt axes pjx, pjy")
        dt = np.inf
        for j in range(v.shape[0]):   # y slices: the rotated bound needs several temporaries per cell
            dt = min(dt, _tti_dt_slice(v[j], np.asarray(eps[j], np.float64), np.asarray(dl[j], np.float64),
                                       np.asarray(pjx[j], np.float32), np.asarray(pjy[j], np.float32),
                                       1.0 / dx, 1.0 / dy, 1.0 / dz, k1max))
        return float(dt)
    e = np.asarray(eps, np.float64)
    dd = np.asarray(dl, np.float64)
    tmp = (1 + 2 * e) * (rdx2 + rdy2) + rdz2
    lam = v * np.sqrt((tmp + np.sqrt(np.maximum(tmp * tmp - 8 * (e - dd) * (rdx2 + rdy2) * rdz2, 0))) / 2)
    return float(np.min(2.0 / k1max / lam / np.sqrt(1 + 2 * e)))


def _tti_dt_slice(v, e, dd, pjx, pjy, akx, aky, akz, k1max):
    """The TTI branch of bgprtm/ModelPrepare.cpp:826-870 on one slice: sin/cos of tilt and azimuth from the
    axis (pjx, pjy), the rotated squared components, and the max over the axis subsets ic = 1..7."""
    st = np.sqrt(pjx.astype(np.float64) ** 2 + pjy.astype(np.float64) ** 2)
    ct = np.sqrt(np.maximum(1.0 - st * st, 0.0))
    big = np.abs(st) > 1e-15
    safe = np.where(big, st, 1.0)
    cp = np.where(big, -pjx / safe, 1.0)
    sp = np.where(big, -pjy / safe, 0.0)
    rx = (cp * ct * akx + sp * ct * aky - st * akz) ** 2
    ry = (-sp * akx + cp * aky) ** 2
    rz = (cp * st * akx + sp * st * aky + ct * akz) ** 2
    lam = np.zeros_like(v)
    for ic in range(1, 8):
        icz, icx, icy = ic & 1, (ic >> 1) & 1, (ic >> 2) & 1
        h = rx * icx + ry * icy
        t1 = (1.0 + 2.0 * e) * h + rz * icz
        t2 = v * np.sqrt((t1 + np.sqrt(np.maximum(t1 * t1 - 8.0 * (e - dd) * h * rz * icz, 0.0))) / 2.0)
        lam = np.maximum(lam, t2)
    return float(np.min(2.0 / k1max / lam / np.sqrt(1.0 + 2.0 * e)))


def surface_receivers(nx, ny, iz, step=1):
    """Receivers on every `step`-th (x, y) cell at depth index iz (grid units)."""
    ys, xs = np.meshgrid(np.arange(0, ny, step), np.arange(0, nx, step), indexing="ij")
    return np.stack([xs.ravel(), ys.ravel(), np.full(xs.size, float(iz))], axis=1).astype(np.float32)


def ricker(nt, dt, fp):
    t = np.arange(nt) * dt - 1.5 / fp
    a = (np.pi * fp * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


class SaltLayeredModel:
    """Layered model with a salt-like body, evaluated on sub-volumes (apertures) on demand.

    Six flat layers with seeded interface depths and velocities (1500-4000 m/s, increasing with depth
    on average), plus an ellipsoidal body at `vsalt` whose semi-axes and centre are seeded and whose
    boundary is made irregular by a low-order random radial perturbation.  The full model
    (nyF, nxF, nz) is never formed: __call__(x0, y0, nx, ny) returns the (ny, nx, nz) block at
    (x0, y0)."""

    def __init__(self, nxF=1000, nyF=1000, nz=400, d=12.5, seed=505, vsalt=4500.0):
        rng = np.random.default_rng(seed)
        self.nxF, self.nyF, self.nz, self.d = nxF, nyF, nz, d
        depths = np.sort(rng.uniform(0.05, 0.95, 5)) * nz
        vels = np.sort(rng.uniform(1500.0, 4000.0, 6))
        vels[1:] = np.maximum(vels[1:], vels[:-1])
        self.profile = vels[np.searchsorted(depths, np.arange(nz) + 0.5)].astype(np.float32)
        self.c = np.array([rng.uniform(0.3, 0.7) * nyF, rng.uniform(0.3, 0.7) * nxF, rng.uniform(0.45, 0.7) * nz])
        self.ax = np.array([rng.uniform(0.12, 0.25) * nyF, rng.uniform(0.12, 0.25) * nxF, rng.uniform(0.1, 0.2) * nz])
        self.pert = rng.uniform(-0.15, 0.15, (3, 2))   # radius modulation: cos/sin of azimuth orders 1..3
        self.vsalt = np.float32(vsalt)

    def __call__(self, x0, y0, nx, ny):
        y = (np.arange(y0, y0 + ny, dtype=np.float64) - self.c[0]) / self.ax[0]
        x = (np.arange(x0, x0 + nx, dtype=np.float64) - self.c[1]) / self.ax[1]
        z = (np.arange(self.nz, dtype=np.float64) - self.c[2]) / self.ax[2]
        v = np.broadcast_to(self.profile[None, None, :], (ny, nx, self.nz)).copy()
        phi = np.arctan2(y[:, None], x[None, :])
        rr = 1.0 + sum(self.pert[m, 0] * np.cos((m + 1) * phi) + self.pert[m, 1] * np.sin((m + 1) * phi)
                       for m in range(3))
        h2 = (y[:, None] ** 2 + x[None, :] ** 2) / (rr * rr)          # (ny, nx)
        inside = (h2[:, :, None] + z[None, None, :] ** 2) <= 1.0
        v[inside] = self.vsalt
        return v
