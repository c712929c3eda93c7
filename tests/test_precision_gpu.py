"""Image accuracy of reduced-precision snapshot storage against an FP64 oracle.

The GPU shot (FP32 stencil; snapshots stored as FP32, FP16 or BF16 with one FP32
scale per 8^3 block, or the reference's COMPRESS_16 code) is compared with the
same shot computed entirely in FP64 by the oracle: NRMSE = ||img - img64|| /
||img64||.  Also the FP64 multi-shot stack.  Small TTI FD16 and VTI CFD10
grids; the measured NRMSE values are printed for the record.
"""
import numpy as np
import pytest

from oracle import pipeline
from konglong_amd._lib import ENGINE_FD, ENGINE_CFD, VTI, TTI, SNAP_FP32, SNAP_C16, SNAP_FP16, SNAP_BF16
from konglong_amd.models import cfl_dt
from konglong_amd.propagator import Propagator, shot_config
from tests.helpers import smooth_field, tilt_axis

pytestmark = pytest.mark.gpu

N = 44
D = 12.5
NB = 10


def nrmse(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _case(engine, order, model, seed=31):
    shape = (N, N, N)
    vel = smooth_field(shape, 1600.0, 2800.0, seed)
    eps = smooth_field(shape, 0.0, 0.25, seed + 1)
    dl = smooth_field(shape, 0.0, 0.12, seed + 2)
    pjx = pjy = None
    if model == TTI:
        pjx, pjy = tilt_axis(shape, seed + 3)
    dt = 0.9 * cfl_dt(vel, D, engine, order, VTI, eps, dl)
    nt = int(0.36 / dt)
    rcv = np.array([[x, y, NB + 2.0] for y in range(NB, N - NB, 3) for x in range(NB, N - NB, 3)], np.float32)
    return vel, eps, dl, pjx, pjy, dt, nt, rcv


def _gpu_shot(engine, order, model, vel, eps, dl, pjx, pjy, dt, nt, rcv, src, fmt, wavelet):
    cfg = shot_config(N, N, N, D, dt, nt, engine, order, model, nbnd=NB, vmax=float(vel.max()), src_xyz=src,
                      f_peak=18.0, spread=5, khpass=0.3, khcut=0.3, nrcv=rcv.shape[0], snap_every=3, snap_format=fmt,
                      tmute_frac=0.1)
    p = Propagator(cfg)
    p.set_model(vel, eps, dl, pjx, pjy)
    p.set_receivers(rcv)
    p.set_source_wavelet(wavelet)
    rec = p.model()
    out = p.rtm(rec)
    p.close()
    return out


def _oracle64(engine, order, model, vel, eps, dl, pjx, pjy, dt, nt, rcv, src):
    o = pipeline.Shot64(N, N, N, D, dt, nt, engine, order, model, [NB] * 6, float(vel.max()), src, 18.0, 5, 0.3, 0.3,
                        rcv, snap_every=3, tmute_frac=0.1, threads=8)
    o.set_model(vel, eps, dl, pjx, pjy)
    img, _ = o.rtm(o.model())
    return img, o.wavelet


@pytest.mark.parametrize("case", [(ENGINE_FD, 16, TTI), (ENGINE_CFD, 10, VTI)], ids=["TTI-FD16", "VTI-CFD10"])
def test_snapshot_precision_nrmse(klrtm, case):
    engine, order, model = case
    vel, eps, dl, pjx, pjy, dt, nt, rcv = _case(engine, order, model)
    src = (N / 2, N / 2, NB + 2.0)
    img64, wavelet = _oracle64(engine, order, model, vel, eps, dl, pjx, pjy, dt, nt, rcv, src)
    limits = {SNAP_FP32: 1e-3, SNAP_FP16: 1e-2, SNAP_BF16: 5e-2, SNAP_C16: 1e-2}
    names = {SNAP_FP32: "fp32", SNAP_FP16: "fp16+scale", SNAP_BF16: "bf16+scale", SNAP_C16: "compress16"}
    res = {}
    for fmt in (SNAP_FP32, SNAP_FP16, SNAP_BF16, SNAP_C16):
        out = _gpu_shot(engine, order, model, vel, eps, dl, pjx, pjy, dt, nt, rcv, src, fmt,
                        wavelet.astype(np.float32))
        res[fmt] = (nrmse(out["image"], img64), nrmse(out["image64"], img64))
        print(f"[precision] {'TTI-FD16' if model == TTI else 'VTI-CFD10'} {N}^3 nt={nt} snapshots={names[fmt]}: "
              f"image NRMSE vs FP64 oracle = {res[fmt][0]:.3e} (FP64-accumulated image {res[fmt][1]:.3e})",
              flush=True)
    for fmt, (e32, e64) in res.items():
        assert e32 <= limits[fmt] and e64 <= limits[fmt], (names[fmt], e32, e64)
    # reduced storage costs accuracy in the expected order
    assert res[SNAP_FP32][0] <= res[SNAP_FP16][0] <= res[SNAP_BF16][0]


def test_fp64_multishot_stack(klrtm):
    """Three shots stacked in FP64 on the GPU against the FP64 oracle's stack (TTI FD16, FP16 snapshots)."""
    engine, order, model = ENGINE_FD, 16, TTI
    vel, eps, dl, pjx, pjy, dt, nt, rcv = _case(engine, order, model, seed=41)
    stack_gpu = np.zeros((N, N, N), np.float64)
    stack_64 = np.zeros((N, N, N), np.float64)
    for sx in (N / 3, N / 2, 2 * N / 3):
        src = (sx, N / 2, NB + 2.0)
        img64, wavelet = _oracle64(engine, order, model, vel, eps, dl, pjx, pjy, dt, nt, rcv, src)
        out = _gpu_shot(engine, order, model, vel, eps, dl, pjx, pjy, dt, nt, rcv, src, SNAP_FP16,
                        wavelet.astype(np.float32))
        stack_gpu += out["image64"]
        stack_64 += img64
    e = nrmse(stack_gpu, stack_64)
    print(f"[precision] TTI-FD16 3-shot FP64 stack, FP16 snapshots: NRMSE vs FP64 oracle = {e:.3e}", flush=True)
    assert e <= 1e-2


@pytest.mark.parametrize("case", [(ENGINE_FD, 16, TTI), (ENGINE_CFD, 10, VTI)], ids=["TTI-FD16", "VTI-CFD10"])
def test_learned_snapshot_nrmse(klrtm, case):
    """Learned-codec snapshots (32 BF16 code values + int8 residual per 8^3 block, trained on ISO
    forward snapshots of other random models on the same grid) against the FP64 oracle, beside
    FP32 snapshots."""
    from konglong_amd import learned
    from konglong_amd._lib import SNAP_LEARNED
    engine, order, model = case
    n = 48
    shape = (n, n, n)
    vel = smooth_field(shape, 1600.0, 2800.0, 31)
    eps = smooth_field(shape, 0.0, 0.25, 32)
    dl = smooth_field(shape, 0.0, 0.12, 33)
    pjx = pjy = None
    if model == TTI:
        pjx, pjy = tilt_axis(shape, 34)
    dt = 0.9 * cfl_dt(vel, D, engine, order, VTI, eps, dl)
    nt = int(0.36 / dt)
    rcv = np.array([[x, y, NB + 2.0] for y in range(NB, n - NB, 3) for x in range(NB, n - NB, 3)], np.float32)
    src = (n / 2, n / 2, NB + 2.0)
    codec, info = learned.train(shape, seeds=[501, 502, 503, 504], d=D, dt=dt, nt=nt, snap_every=3, nl=32, nbnd=NB,
                                f_peak=18.0, spread=5, vmin=1600.0, vmax=2800.0, sigma=4.0)
    print(f"[precision] learned codec trained on {info['snapshots']} snapshots, energy kept {info['energy_kept']:.4f}",
          flush=True)
    o = pipeline.Shot64(n, n, n, D, dt, nt, engine, order, model, [NB] * 6, float(vel.max()), src, 18.0, 5, 0.3, 0.3,
                        rcv, snap_every=3, tmute_frac=0.1, threads=8)
    o.set_model(vel, eps, dl, pjx, pjy)
    img64, _ = o.rtm(o.model())
    res = {}
    for fmt in (SNAP_FP32, SNAP_LEARNED):
        cfg = shot_config(n, n, n, D, dt, nt, engine, order, model, nbnd=NB, vmax=float(vel.max()), src_xyz=src,
                          f_peak=18.0, spread=5, khpass=0.3, khcut=0.3, nrcv=rcv.shape[0], snap_every=3, snap_format=fmt,
                          tmute_frac=0.1)
        p = Propagator(cfg)
        if fmt == SNAP_LEARNED:
            from konglong_amd._lib import lib as _l, check as _c
            _c(_l().klrtm_propagator_set_lcodec(p.h, codec.h))
        p.set_model(vel, eps, dl, pjx, pjy)
        p.set_receivers(rcv)
        p.set_source_wavelet(o.wavelet.astype(np.float32))
        out = p.rtm(p.model())
        p.close()
        res[fmt] = nrmse(out["image64"], img64)
    name = "TTI-FD16" if model == TTI else "VTI-CFD10"
    ratio = n ** 3 * 4 / codec.code_bytes(shape)
    print(f"[precision] {name} {n}^3 nt={nt}: image NRMSE vs FP64 oracle, fp32 snapshots {res[SNAP_FP32]:.3e}, "
          f"learned ({ratio:.2f}x smaller than fp32) {res[SNAP_LEARNED]:.3e}", flush=True)
    assert res[SNAP_LEARNED] <= 1e-2
    codec.close()
