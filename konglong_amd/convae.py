"""Convolutional autoencoder snapshot codec (klrtm_convae): training on propagator snapshots, encode /
decode of volumes, and its use as the snapshot format of the propagator.

The codec replaces the snapshot compression of bgprtm's cache (WaveFieldCompress,
bgprtm/WaveFieldCompress.cpp:25-279) with 32^3-block convolutions learned from snapshots the propagator
generates on seeded random velocity models; the network and its matrix-core kernels are described in
konglong_amd/csrc/convae.hip."""
import ctypes as C

import numpy as np

from ._lib import DeviceArray, check, lib, require, ENGINE_FD, ISO, SNAP_FP32
from .learned import random_velocity
from .propagator import Propagator, shot_config


class ConvAutoencoder:
    def __init__(self, seed=1, params=None):
        h = C.c_void_p()
        check(lib().klrtm_convae_create(C.byref(h), int(seed)))
        self.h = h.value
        self.n = int(lib().klrtm_convae_param_count())
        if params is not None:
            self.set_params(params)

    def params(self):
        p = np.empty(self.n, np.float32)
        check(lib().klrtm_convae_get_params(self.h, p.ctypes.data))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        require(p.size == self.n, f"the autoencoder has {self.n} parameters, got {p.size}")
        check(lib().klrtm_convae_set_params(self.h, p.ctypes.data))

    @staticmethod
    def code_bytes(shape):
        ny, nx, nz = shape
        return int(lib().klrtm_convae_code_bytes(nx, ny, nz))

    def encode(self, vol, code, floor_frac=0.05):
        ny, nx, nz = vol.shape
        check(lib().klrtm_convae_encode(self.h, vol.ptr, nx, ny, nz, float(floor_frac), code.ptr))

    def decode(self, code, vol):
        ny, nx, nz = vol.shape
        check(lib().klrtm_convae_decode(self.h, code.ptr, vol.ptr, nx, ny, nz))

    def fit(self, bank, steps, batch=64, lr=1e-3, floor_frac=0.05, seed=0):
        """Adam on random 32^3 blocks of a device bank of snapshots (DeviceArray [nsnap][ny][nx][nz]).
        Returns the first and last batch loss."""
        nsnap, ny, nx, nz = bank.shape
        l0, l1 = C.c_float(), C.c_float()
        check(lib().klrtm_convae_train(self.h, bank.ptr, nsnap, nx, ny, nz, float(floor_frac), int(steps), int(batch),
                                       float(lr), int(seed), C.byref(l0), C.byref(l1)))
        return l0.value, l1.value

    def save(self, path):
        np.save(path, self.params())

    @classmethod
    def load(cls, path):
        return cls(params=np.load(path, allow_pickle=False))

    def close(self):
        if getattr(self, "h", None):
            lib().klrtm_convae_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def snapshot_bank(shape, seeds, d, dt, nt, per_model=4, engine=ENGINE_FD, order=8, model=ISO, nbnd=20, f_peak=15.0,
                  vmin=1500.0, vmax=4500.0, sigma=6.0, spread=10):
    """Device bank [len(seeds)*per_model][ny][nx][nz] of forward snapshots: one shot per seeded random
    velocity model (random source position), `per_model` snapshots evenly spread over the second half
    to the end of nt steps, when the wavefield fills the model."""
    ny, nx, nz = shape
    n = ny * nx * nz
    bank = DeviceArray((len(seeds) * per_model, ny, nx, nz))
    times = np.linspace(nt // 2, nt - 1, per_model).astype(int)
    w0, w1 = C.c_void_p(), C.c_void_p()
    k = 0
    for seed in seeds:
        vel = random_velocity(shape, seed, vmin, vmax, sigma)
        rng = np.random.default_rng(seed + 7919)
        src = (float(rng.uniform(nbnd, nx - 1 - nbnd)), float(rng.uniform(nbnd, ny - 1 - nbnd)),
               float(rng.uniform(nbnd + 2, nz / 2)))
        cfg = shot_config(nx, ny, nz, d, dt, nt, engine, order, model, nbnd=nbnd, vmax=float(vel.max()), src_xyz=src,
                          f_peak=f_peak, spread=spread, snap_every=1, snap_format=SNAP_FP32)
        p = Propagator(cfg)
        p.set_model(vel)
        p.reset()
        it = 0
        for t in times:
            p.forward_steps(it, t + 1 - it, False)
            it = t + 1
            check(lib().klrtm_propagator_wavefields(p.h, C.byref(w0), C.byref(w1)))
            check(lib().klrtm_memcpy_d2d(bank.ptr + k * n * 4, w1.value, n * 4))
            k += 1
        p.close()
    return bank


def train(shape, seeds, d, dt, nt, steps, batch=64, lr=1e-3, per_model=4, floor_frac=0.05, seed=1, **kw):
    """A codec trained on snapshots of one shot per seeded model; returns (ConvAutoencoder, info)."""
    bank = snapshot_bank(shape, seeds, d, dt, nt, per_model, **kw)
    net = ConvAutoencoder(seed)
    l0, l1 = net.fit(bank, steps, batch, lr, floor_frac, seed)
    info = dict(snapshots=bank.shape[0], steps=steps, batch=batch, loss_first=l0, loss_last=l1)
    bank.free()
    return net, info


def _bench(n=400, train_steps=20, reps=3):
    """Times encode and decode of one n^3 snapshot (smooth random field, untrained weights) and
    `train_steps` Adam steps on 64-block batches of a small random bank; used for the codec's kernel
    profile."""
    import time
    from ._lib import synchronize
    rng = np.random.default_rng(0)
    net = ConvAutoencoder(5)
    bank = DeviceArray.from_host(rng.standard_normal((2, 96, 96, 96)).astype(np.float32))
    synchronize()
    t0 = time.perf_counter()
    net.fit(bank, train_steps, 64, 1e-3)
    synchronize()
    res = {"train_ms_per_step": (time.perf_counter() - t0) / train_steps * 1e3}
    bank.free()
    shape = (n, n, n)
    vol = DeviceArray.from_host(rng.standard_normal(shape).astype(np.float32))
    code = DeviceArray((ConvAutoencoder.code_bytes(shape),), np.uint8)
    out = DeviceArray(shape)
    for name, f in (("encode_ms", lambda: net.encode(vol, code)), ("decode_ms", lambda: net.decode(code, out))):
        f()
        synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        synchronize()
        res[name] = (time.perf_counter() - t0) / reps * 1e3
    net.close()
    return res


if __name__ == "__main__":
    import json
    import sys
    print(json.dumps({k: round(v, 4) for k, v in _bench(int(sys.argv[1]) if len(sys.argv) > 1 else 400).items()}))
