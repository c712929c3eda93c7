"""Parity at the workloads' full sizes.

* ISO FD8 RTM shot on 128^3 (dx 10 m, dt 1 ms, PML 20, 600 steps, layered model with a Gaussian anomaly,
  15 Hz source at the centre, a receiver on every surface cell 2 cells below the top PML, spread 10, all
  600 snapshots in FP32): record and image against the FP32 oracle shot, 1e-4 relative L2.
* CFD10 VTI steps on 400^3 (dx 12.5 m, PML 32, the smooth random VTI model): the fused step and the
  KernelCPU::VTI accumulate bit for bit against the oracle.
* TTI FD16 steps on 600x600x400 (dx 12.5 m, PML 40, smooth tilt and azimuth): 1e-4 relative L2 (the FD
  kernels contract multiply-adds).
The oracle runs on the host's cores (OpenMP)."""
import os

import numpy as np
import pytest

import oracle
from oracle import pipeline
from konglong_amd import models
from konglong_amd._lib import DeviceArray, ENGINE_FD, ENGINE_CFD, ISO, VTI, TTI, SNAP_FP32
from konglong_amd.kernel import KernelCPU
from konglong_amd.propagator import Propagator, shot_config
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def test_iso_fd8_128_full_shot(klrtm):
    n, d, dt, nt, nb = 128, 10.0, 1e-3, 600, 20
    print(f"[config-size] ISO FD8 {n}^3 nt={nt}, oracle threads {THREADS}", flush=True)
    vel = models.layered_model(n, d, 101)
    iz = nb + 2
    rcv = models.surface_receivers(n, n, iz)
    src = (64.0, 64.0, 64.0)
    cfg = shot_config(n, n, n, d, dt, nt, ENGINE_FD, 8, ISO, nbnd=nb, vmax=float(vel.max()), src_xyz=src, f_peak=15.0,
                      spread=10, nrcv=rcv.shape[0], snap_every=1, snap_format=SNAP_FP32, tmute_frac=0.1)
    p = Propagator(cfg)
    p.set_model(vel)
    p.set_receivers(rcv)
    o = pipeline.Shot(n, n, n, d, dt, nt, ENGINE_FD, 8, ISO, [nb] * 6, float(vel.max()), src, 15.0, 10, 0.3, 0.3, rcv,
                      snap_every=1, tmute_frac=0.1, threads=THREADS)
    o.set_model(vel)
    p.set_source_wavelet(o.wavelet)
    want = o.model()
    rec = p.model()
    err = rel_l2(rec, want)
    print(f"  record rel_l2={err:.2e} (|rec|max {np.abs(want).max():.3e})", flush=True)
    assert np.abs(want).max() > 0 and err < 1e-4
    img_want, ill_want = o.rtm(want)
    out = p.rtm(want, want=("image", "illum"))
    e_img, e_ill = rel_l2(out["image"], img_want), rel_l2(out["illum"], ill_want)
    print(f"  image rel_l2={e_img:.2e} illum rel_l2={e_ill:.2e}", flush=True)
    assert np.abs(img_want).max() > 0 and e_img < 1e-4 and e_ill < 1e-4
    p.close()


def _engine_pair(shape, d, engine, order, model, nb, seed, pj_seed=None):
    ny, nx, nz = shape
    vel, eps, dl = models.smooth_vti_model(shape, 20.0, seed)
    pjx = pjy = None
    if model == TTI:
        pjx, pjy = models.tilt_axes(shape, 20.0, pj_seed)
    dt = models.cfl_dt(vel, d, engine, order, model, eps, dl, pjx, pjy)
    m = (vel * vel * (np.float32(dt) * np.float32(dt))).astype(np.float32)
    vmax = float(vel.max())
    k = KernelCPU(nx, ny, nz, d, d, d, engine, order)
    dev = {name: DeviceArray.from_host(a) for name, a in
           (("m", m), ("eps", eps), ("del", dl), ("pjx", pjx), ("pjy", pjy)) if a is not None}
    k.setModel(dev["m"], None, dev["del"], dev["eps"], dev.get("pjx"), dev.get("pjy"))
    k.setBoundary(*([nb] * 6), dt, vmax)
    o = oracle.Engine(nx, ny, nz, d, d, d, engine, order, np.float32, threads=THREADS)
    o.set_model(m, eps, dl, pjx, pjy)
    o.set_boundary([nb] * 6, dt, vmax, dim_pml=3)
    # a smooth pulse plus noise, so every cell and the PML layers carry signal
    rng = np.random.default_rng(seed + 1)
    w1 = (models.gaussian_smooth_noise(shape, 6.0, seed + 2) + 1e-2 * rng.standard_normal(shape)).astype(np.float32)
    w0 = (0.97 * w1).astype(np.float32)
    return k, dev, o, w1, w0


def test_vti_cfd10_400_steps_bit_exact(klrtm):
    shape, nb, steps = (400, 400, 400), 32, 3
    print(f"[config-size] VTI CFD10 {shape} {steps} steps, oracle threads {THREADS}", flush=True)
    k, dev, o, w1, w0 = _engine_pair(shape, 12.5, ENGINE_CFD, 10, VTI, nb, 202)
    d1, d0, dr = DeviceArray.from_host(w1), DeviceArray.from_host(w0), DeviceArray(shape)
    ow1, ow0, owr = w1.copy(), w0.copy(), np.zeros_like(w1)
    for it in range(steps):   # fused step (kernel + leapfrog), the bench's form
        k.step(VTI, d1, d0, dr)
        o.step(VTI, ow1, ow0, owr)
        d0, d1 = d1, d0
        ow0, ow1 = ow1, ow0
        print(f"  step {it} done", flush=True)
    got = d1.download()
    same = np.array_equal(got.view(np.uint32), ow1.view(np.uint32))
    print(f"  fused steps: bit-exact={same} rel_l2={rel_l2(got, ow1):.2e}", flush=True)
    assert same
    # KernelCPU::VTI (three accumulating passes, pr an output) from the current state
    pb_dev = DeviceArray.from_host(ow0)
    pb = ow0.copy()
    k.VTI(d1, pb_dev, dr)
    o.kernel(VTI, ow1, pb, owr)
    assert np.array_equal(pb_dev.download().view(np.uint32), pb.view(np.uint32))
    o.close()
    k.close()


def test_tti_fd16_600_steps(klrtm):
    shape, nb, steps = (600, 600, 400), 40, 3
    print(f"[config-size] TTI FD16 {shape} {steps} steps, oracle threads {THREADS}", flush=True)
    k, dev, o, w1, w0 = _engine_pair(shape, 12.5, ENGINE_FD, 16, TTI, nb, 303, 313)
    d1, d0, dr = DeviceArray.from_host(w1), DeviceArray.from_host(w0), DeviceArray(shape)
    ow1, ow0, owr = w1.copy(), w0.copy(), np.zeros_like(w1)
    for it in range(steps):
        k.step(TTI, d1, d0, dr)
        o.step(TTI, ow1, ow0, owr)
        d0, d1 = d1, d0
        ow0, ow1 = ow1, ow0
        print(f"  step {it} done", flush=True)
    err = rel_l2(d1.download(), ow1)
    print(f"  rel_l2={err:.2e}", flush=True)
    assert err < 1e-4
    o.close()
    k.close()
