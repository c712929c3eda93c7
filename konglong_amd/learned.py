"""Learned snapshot codec: a linear autoencoder on 8x8x8 wavefield blocks.

It takes the place of WaveFieldCompress in the source-wavefield snapshot cache
(bgprtm/WaveFieldCompress.cpp:25-279, called from bgprtm/Propagator.cpp:
1233-1243).  A block's 512 samples x are stored as nl BF16 values z = x W (the
code of a linear autoencoder) plus the residual x - z W^T quantised to int8
with one scale per block, and decoded as z W^T + q*scale; encode, decode (fused
into the imaging condition) and the training statistics run as BF16
matrix-core kernels with FP32 accumulation (``klrtm_lcodec_*``).

Training uses snapshots the propagator itself generates: forward runs on
seeded random velocity models, the Gram matrix G = sum x^T x of their blocks
accumulated on the device in FP64, and W = the nl leading eigenvectors of G --
the linear code of width nl with the least squared reconstruction error over
the training blocks (the optimum a linear autoencoder trained on that loss
converges to).
"""
import ctypes as C

import numpy as np

from ._lib import DeviceArray, check, lib, require, ENGINE_FD, ISO, SNAP_FP32
from .propagator import Propagator, shot_config

BLOCK = 8
FEATURES = BLOCK ** 3


def block_features(vol):
    """(nblk, 512) matrix of a (ny, nx, nz) volume, blocks in the codec's order
    (b = (by*nbx + bx)*nbz + bz, features (ly*8 + lx)*8 + lz)."""
    ny, nx, nz = vol.shape
    v = vol.reshape(ny // 8, 8, nx // 8, 8, nz // 8, 8)        # by ly bx lx bz lz
    return v.transpose(0, 2, 4, 1, 3, 5).reshape(-1, FEATURES)  # (by bx bz) (ly lx lz)


def from_block_features(X, shape):
    ny, nx, nz = shape
    v = X.reshape(ny // 8, nx // 8, nz // 8, 8, 8, 8).transpose(0, 3, 1, 4, 2, 5)
    return np.ascontiguousarray(v.reshape(ny, nx, nz))


def leading_eigvecs(G, nl):
    """W (512, nl) float32: orthonormal eigenvectors of the symmetric G for its nl largest eigenvalues,
    each signed so that its largest-magnitude entry is positive (a deterministic choice)."""
    G = np.asarray(G, np.float64)
    G = 0.5 * (G + G.T)
    w, V = np.linalg.eigh(G)
    V = V[:, ::-1][:, :nl]
    s = np.sign(V[np.abs(V).argmax(axis=0), np.arange(nl)])
    s[s == 0] = 1.0
    return np.ascontiguousarray((V * s).astype(np.float32)), w[::-1]


class GramAccumulator:
    """Device FP64 Gram matrix of 8x8x8 blocks, G += sum_b (s x_b)^T (s x_b)."""

    def __init__(self):
        self.G = DeviceArray((FEATURES, FEATURES), np.float64)
        self.count = 0

    def add(self, vol_ptr, shape, scale=1.0):
        ny, nx, nz = shape
        check(lib().klrtm_lcodec_gram(vol_ptr, nx, ny, nz, float(scale), self.G.ptr))
        self.count += (nx // 8) * (ny // 8) * (nz // 8)

    def matrix(self):
        return self.G.download()


class LearnedCompressor:
    """klrtm_lcodec with W (512, nl), nl = 32, 64, 128 or 256 latent values per 512-sample block."""

    def __init__(self, W):
        W = np.ascontiguousarray(W, np.float32)
        require(W.shape[0] == FEATURES and W.shape[1] in (32, 64, 128, 256), f"code matrix of shape {W.shape}")
        self.W = W
        self.nl = W.shape[1]
        h = C.c_void_p()
        check(lib().klrtm_lcodec_create(C.byref(h), self.nl, W.ctypes.data_as(C.POINTER(C.c_float))))
        self.h = h.value

    def code_bytes(self, shape):
        ny, nx, nz = shape
        return int(lib().klrtm_lcodec_code_bytes(self.h, nx, ny, nz))

    def new_code(self, shape):
        return DeviceArray((self.code_bytes(shape),), np.uint8)

    def split_code(self, code_host, shape):
        """(z BF16 bits [nblk][nl], scale [nblk], q int8 [nblk][512]) views of a downloaded code."""
        ny, nx, nz = shape
        nblk = (ny // 8) * (nx // 8) * (nz // 8)
        a256 = lambda v: (v + 255) // 256 * 256
        zs, ss = a256(nblk * self.nl * 2), a256(nblk * 4)
        z = code_host[:nblk * self.nl * 2].view(np.uint16).reshape(nblk, self.nl)
        sc = code_host[zs:zs + nblk * 4].view(np.float32)
        q = code_host[zs + ss:zs + ss + nblk * 512].view(np.int8).reshape(nblk, 512)
        return z, sc, q

    def encode(self, vol, code):
        """vol: DeviceArray (ny, nx, nz) float32 -> code: DeviceArray of code_bytes bytes."""
        ny, nx, nz = vol.shape
        check(lib().klrtm_lcodec_encode(self.h, vol.ptr, code.ptr, nx, ny, nz))

    def decode(self, code, vol):
        ny, nx, nz = vol.shape
        check(lib().klrtm_lcodec_decode(self.h, code.ptr, vol.ptr, nx, ny, nz))

    def close(self):
        if getattr(self, "h", None):
            lib().klrtm_lcodec_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def save(self, path):
        np.save(path, self.W)

    @classmethod
    def load(cls, path):
        return cls(np.load(path, allow_pickle=False))


def random_velocity(shape, seed, vmin=1500.0, vmax=4500.0, sigma=6.0):
    """Seeded smooth random velocity: white noise, Gaussian filter (sigma cells, FFT), mapped to
    [vmin, vmax]."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(shape).astype(np.float32)
    F = np.fft.rfftn(v)
    k = [np.fft.fftfreq(n) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1])]
    g = np.ones(F.shape, np.float32)
    for ax, kk in enumerate(k):
        sh = [1] * len(shape)
        sh[ax] = -1
        g = g * np.exp(-2.0 * (np.pi * sigma * kk.reshape(sh)) ** 2).astype(np.float32)
    v = np.fft.irfftn(F * g, s=shape, axes=tuple(range(len(shape)))).astype(np.float32)
    lo, hi = float(v.min()), float(v.max())
    return (vmin + (v - lo) * ((vmax - vmin) / max(hi - lo, 1e-30))).astype(np.float32)


def train(shape, seeds, d, dt, nt, snap_every, nl=32, engine=ENGINE_FD, order=8, model=ISO, nbnd=20, f_peak=15.0,
          vmin=1500.0, vmax=4500.0, sigma=6.0, spread=10, models=None, eps=None, del_=None, normalize=True):
    """Train the codec on forward snapshots of one shot per seeded random model.

    Every snap_every-th step of each run is added to the Gram matrix (scaled to unit RMS when
    normalize, so that every snapshot weighs the same).  Returns (LearnedCompressor, info)."""
    ny, nx, nz = shape
    acc = GramAccumulator()
    nsnap = 0
    for i, seed in enumerate(seeds):
        vel = models[i] if models is not None else random_velocity(shape, seed, vmin, vmax, sigma)
        rng = np.random.default_rng(seed + 7919)
        src = (float(rng.uniform(nbnd, nx - 1 - nbnd)), float(rng.uniform(nbnd, ny - 1 - nbnd)),
               float(rng.uniform(nbnd + 2, nz / 2)))
        cfg = shot_config(nx, ny, nz, d, dt, nt, engine, order, model, nbnd=nbnd, vmax=float(vel.max()), src_xyz=src,
                          f_peak=f_peak, spread=spread, snap_every=snap_every, snap_format=SNAP_FP32)
        p = Propagator(cfg)
        p.set_model(vel, eps, del_)
        p.reset()
        w0, w1 = C.c_void_p(), C.c_void_p()
        for it0 in range(0, nt, snap_every):
            p.forward_steps(it0, min(snap_every, nt - it0), False)
            check(lib().klrtm_propagator_wavefields(p.h, C.byref(w0), C.byref(w1)))
            s = 1.0
            if normalize:
                host = np.empty(shape, np.float32)
                check(lib().klrtm_memcpy_d2h(host.ctypes.data, w1.value, host.nbytes))
                rms = float(np.sqrt(np.mean(host.astype(np.float64) ** 2)))
                if rms <= 0.0:
                    continue
                s = 1.0 / rms
            acc.add(w1.value, shape, s)
            nsnap += 1
        p.close()
    G = acc.matrix()
    W, ev = leading_eigvecs(G, nl)
    kept = float(ev[:nl].sum() / max(ev.sum(), 1e-300))
    return LearnedCompressor(W), dict(snapshots=nsnap, blocks=acc.count, energy_kept=kept)


def _bench(n=400, nl=32, reps=5):
    """Times encode, decode and decode+imaging of one n^3 snapshot (random W, random field) against the
    FP32 imaging kernel; used for the codec's kernel profile."""
    import time
    from ._lib import synchronize
    rng = np.random.default_rng(0)
    W, _ = np.linalg.qr(rng.standard_normal((FEATURES, nl)))
    c = LearnedCompressor(W.astype(np.float32))
    shape = (n, n, n)
    x = DeviceArray.from_host(rng.standard_normal(shape).astype(np.float32))
    w = DeviceArray.from_host(rng.standard_normal(shape).astype(np.float32))
    img, ill, out = DeviceArray(shape), DeviceArray(shape), DeviceArray(shape)
    code = c.new_code(shape)
    L = lib()
    res = {}
    for name, f in (("encode", lambda: c.encode(x, code)), ("decode", lambda: c.decode(code, out)),
                    ("decode+imaging", lambda: check(L.klrtm_lcodec_correlate(c.h, code.ptr, w.ptr, n, n, n, 0.5, None,
                                                                              0, img.ptr, ill.ptr, None, None))),
                    ("fp32 imaging", lambda: check(L.klrtm_ic_correlate(x.ptr, w.ptr, n, n, n, 0.5, None, 0, img.ptr,
                                                                        ill.ptr, None, None)))):
        f()
        synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        synchronize()
        res[name] = (time.perf_counter() - t0) / reps * 1e3
    c.close()
    return res


if __name__ == "__main__":
    import json
    import sys
    print(json.dumps({k: round(v, 4) for k, v in _bench(int(sys.argv[1]) if len(sys.argv) > 1 else 400).items()}))
