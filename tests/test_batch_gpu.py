"""Batched multi-shot RTM (BatchRTM) on uneven apertures.

Three ISO FD8 shots on their own aperture sub-grids of one output grid (overlapping, different sizes)
are migrated three ways: the plain path (one shot at a time, FP32 snapshots, launches issued one by
one), a batch of all three side by side with each shot's time loop as captured-graph replays (FP32
snapshots), and the same batch with learned snapshots.  The FP32 batched stack equals the plain stack
bit for bit (the same per-shot launches, stacked in shot order).  The learned stack is measured
against the same three shots computed in FP64 by the oracle and stacked in FP64: NRMSE <= 1e-2, the
bound the learned path is held to."""
import numpy as np
import pytest

from oracle import pipeline
from konglong_amd import learned
from konglong_amd._lib import ENGINE_FD, ISO, SNAP_FP32, SNAP_LEARNED
from konglong_amd.batch import BatchRTM, Shot
from tests.helpers import smooth_field

pytestmark = pytest.mark.gpu

NYF, NXF, NZ = 72, 80, 40
D, DT, NT, NB, K = 10.0, 1.0e-3, 150, 8, 3
APERTURES = [(0, 0, 40, 32), (24, 16, 48, 40), (40, 24, 32, 48)]  # x0, y0, nx, ny


def _shots():
    shots = []
    for x0, y0, nx, ny in APERTURES:
        rcv = np.array([[x, y, NB + 2.0] for y in range(NB, ny - NB, 2) for x in range(NB, nx - NB, 2)], np.float32)
        shots.append(Shot(x0, y0, nx, ny, ((nx - 1) / 2, (ny - 1) / 2, NB + 2.0), rcv))
    return shots


def nrmse(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_batched_shots_match_plain_and_fp64_oracle(klrtm):
    print(f"[batch] {len(APERTURES)} shots on {NXF}x{NYF}x{NZ}, nt={NT}, snap_every={K}", flush=True)
    vfull = smooth_field((NYF, NXF, NZ), 1500.0, 2500.0, 77)

    def vel(x0, y0, nx, ny):
        return np.ascontiguousarray(vfull[y0:y0 + ny, x0:x0 + nx, :])

    common = dict(d=D, dt=DT, nt=NT, engine=ENGINE_FD, order=8, model=ISO, nbnd=NB, f_peak=20.0, spread=5,
                  snap_every=K, tmute_frac=0.1, vmax=2500.0)
    shots = _shots()
    plain = BatchRTM((NYF, NXF, NZ), vel, common, batch=1, snap_format=SNAP_FP32)
    records = plain.run(shots)
    s_plain = plain.stack.download()
    fast32 = BatchRTM((NYF, NXF, NZ), vel, common, batch=3, snap_format=SNAP_FP32, graph=True)
    fast32.run(shots, records)
    assert fast32.times["groups"] == [3]
    assert np.array_equal(fast32.stack.download().view(np.uint64), s_plain.view(np.uint64))
    print("[batch] captured FP32 batch stack == plain stack (bit for bit)", flush=True)

    codec, info = learned.train((40, 40, 40), [901, 902], D, DT, NT, K, nl=32, nbnd=NB, f_peak=20.0, vmin=1500.0,
                                vmax=2500.0, spread=5)
    fastl = BatchRTM((NYF, NXF, NZ), vel, common, batch=3, snap_format=SNAP_LEARNED, lcodec=codec, graph=True)
    fastl.run(shots, records)
    s_learned = fastl.stack.download()
    eagerl = BatchRTM((NYF, NXF, NZ), vel, common, batch=3, snap_format=SNAP_LEARNED, lcodec=codec)
    eagerl.run(shots, records)
    assert np.array_equal(eagerl.stack.download().view(np.uint64), s_learned.view(np.uint64))

    # the FP64 oracle: each shot modelled and migrated in FP64, stacked in FP64
    s64 = np.zeros((NYF, NXF, NZ), np.float64)
    for sh in shots:
        o = pipeline.Shot64(sh.nx, sh.ny, NZ, D, DT, NT, ENGINE_FD, 8, ISO, [NB] * 6, 2500.0, sh.src_xyz, 20.0, 5, 0.3,
                            0.3, sh.rcv_xyz, snap_every=K, tmute_frac=0.1, threads=8)
        o.set_model(vel(sh.x0, sh.y0, sh.nx, sh.ny))
        img, _ = o.rtm(o.model())
        s64[sh.y0:sh.y0 + sh.ny, sh.x0:sh.x0 + sh.nx, :] += img
    e_plain, e_learned = nrmse(s_plain, s64), nrmse(s_learned, s64)
    print(f"[batch] stack NRMSE vs FP64 oracle: plain FP32 {e_plain:.3e}, batched learned {e_learned:.3e} "
          f"(codec energy kept {info['energy_kept']:.6f})", flush=True)
    assert e_plain <= 1e-3
    assert e_learned <= 1e-2
    for b in (plain, fast32, fastl, eagerl, codec):
        b.close()
