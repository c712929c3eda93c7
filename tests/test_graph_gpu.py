"""Captured step sequences: the stepping API with klrtm_propagator_set_graph replays whole
snapshot-interval groups as hipGraphs driven by a device step counter.  Every replay must give the
same bits as the launches issued one by one: the forward wavefield after the source run, the stored
snapshots (through the image they produce) and the image and illumination of the backward run, for
each snapshot format (the encoders write a staging buffer that a counter-indexed copy moves into the
slot) and for odd and even snapshot intervals (odd intervals capture two of them)."""
import ctypes as C

import numpy as np
import pytest

from konglong_amd import models
from konglong_amd._lib import (ENGINE_CFD, ENGINE_FD, ISO, VTI, SNAP_BF16, SNAP_C16, SNAP_CONVAE, SNAP_FP32,
                               SNAP_LEARNED, check, lib)
from konglong_amd.convae import ConvAutoencoder
from konglong_amd.learned import LearnedCompressor
from konglong_amd.propagator import Propagator, shot_config
from tests.helpers import smooth_field

pytestmark = pytest.mark.gpu

SHAPE = (24, 32, 40)  # ny, nx, nz (multiples of 8 for the learned codec)
D, DT, NT, NB = 10.0, 1.0e-3, 61, 8

CASES = [
    # engine, order, model, snapshot format, snapshot interval, receiver wavefield through the codec
    (ENGINE_FD, 8, ISO, SNAP_FP32, 4, 0),
    (ENGINE_FD, 8, ISO, SNAP_LEARNED, 3, 0),
    (ENGINE_FD, 8, ISO, SNAP_BF16, 5, 0),
    (ENGINE_CFD, 10, VTI, SNAP_C16, 2, 1),
    (ENGINE_FD, 8, ISO, SNAP_CONVAE, 4, 0),
]
IDS = ["iso-fp32-k4", "iso-learned-k3", "iso-bf16-k5", "vti-cfd-c16-k2-rq", "iso-convae-k4"]


def _images(p):
    img, ill = C.c_void_p(), C.c_void_p()
    i64, l64 = C.c_void_p(), C.c_void_p()
    check(lib().klrtm_propagator_images(p.h, C.byref(img), C.byref(ill), C.byref(i64), C.byref(l64)))
    n = int(np.prod(SHAPE))
    out = []
    for ptr, dt in ((img, np.float32), (ill, np.float32), (i64, np.float64), (l64, np.float64)):
        a = np.empty(n, dt)
        check(lib().klrtm_memcpy_d2h(a.ctypes.data, ptr.value, a.nbytes))
        out.append(a)
    return out


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_captured_steps_match_eager(klrtm, case):
    engine, order, model, fmt, k, rq = case
    print(f"[graph] {IDS[CASES.index(case)]} shape={SHAPE} nt={NT}", flush=True)
    ny, nx, nz = SHAPE
    vel = smooth_field(SHAPE, 1500.0, 2500.0, 31)
    eps = smooth_field(SHAPE, 0.0, 0.2, 32) if model != ISO else None
    dl = smooth_field(SHAPE, 0.0, 0.1, 33) if model != ISO else None
    rcv = models.surface_receivers(nx, ny, NB + 2, step=2)
    codec = None
    if fmt == SNAP_LEARNED:
        Q, _ = np.linalg.qr(np.random.default_rng(5).standard_normal((512, 32)))
        codec = LearnedCompressor(Q.astype(np.float32))
    elif fmt == SNAP_CONVAE:
        codec = ConvAutoencoder(seed=3)  # untrained weights: the captured and eager runs must still agree
    cfg = shot_config(nx, ny, nz, D, DT, NT, engine, order, model, nbnd=NB, vmax=2500.0,
                      src_xyz=((nx - 1) / 2, (ny - 1) / 2, NB + 2.0), f_peak=25.0, nrcv=rcv.shape[0], snap_every=k,
                      snap_format=fmt, rec_through_codec=rq, tmute_frac=0.2)
    # the record the backward runs inject (modelled once, eagerly)
    p0 = Propagator(cfg)
    p0.set_model(vel, eps, dl)
    p0.set_receivers(rcv)
    record = p0.model()
    p0.close()

    def shot(p):
        p.reset()
        p.forward_steps(0, NT, store_snapshots=True)
        wf = p.wavefield()
        p.reset()
        check(lib().klrtm_propagator_clear_images(p.h))
        p.backward_steps(NT - 1, NT, image=True)
        return [wf] + _images(p)

    runs = []
    for graph in (False, True):
        p = Propagator(cfg)
        if fmt == SNAP_LEARNED:
            check(lib().klrtm_propagator_set_lcodec(p.h, codec.h))
        elif fmt == SNAP_CONVAE:
            check(lib().klrtm_propagator_set_convae(p.h, codec.h, 0.05))
        p.set_model(vel, eps, dl)
        p.set_receivers(rcv)
        p.set_record(record)
        p.set_graph(graph)
        runs.append(shot(p))
        if graph:
            ngraphs, replays = p.graph_stats()
            G = k if k % 2 == 0 else 2 * k
            # forward: NT // G groups from step 0; backward: whole groups below the first step above the cadence
            first = next(it for it in range(NT - 1, -1, -1) if it % k == k - 1)
            assert ngraphs == 2 and replays == NT // G + (first + 1) // G, (ngraphs, replays)
            # a second shot after one eager step on the same stream: the wavefield pointers start swapped,
            # so both directions capture their second graph
            p.forward_steps(0, 1, store_snapshots=False)
            runs.append(shot(p))
            assert p.graph_stats() == (4, 2 * replays)
        p.close()
    if codec is not None:
        codec.close()
    names = ["forward wavefield", "image", "illumination", "image64", "illumination64"]
    assert np.abs(runs[0][1]).max() > 0
    for r in runs[1:]:
        for name, a, b in zip(names, runs[0], r):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
