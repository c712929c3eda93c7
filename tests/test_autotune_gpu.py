"""The tile / block autotuner (konglong_amd/batch.py Autotuner): it times candidates and keeps the fastest per
grid and engine, for the FD stencil tiles and the compact engine's line blocks; a tuned shot records what the
default one records: bit for bit on the compact engine (a line's arithmetic does not depend on the block), to
FP32 rounding on FD (a tile can move a cell between the PML and interior variants of a kernel)."""
import numpy as np
import pytest

from konglong_amd._lib import ENGINE_FD, ENGINE_CFD, ISO, VTI, SNAP_FP32
from konglong_amd.batch import Autotuner
from konglong_amd.propagator import Propagator, shot_config
from tests.helpers import smooth_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [(ENGINE_FD, 8, ISO), (ENGINE_FD, 8, VTI), (ENGINE_CFD, 10, VTI)],
                         ids=["FD8-ISO", "FD8-VTI", "CFD10-VTI"])
def test_autotuner_picks_a_candidate_and_keeps_results(klrtm, case):
    engine, order, model = case
    ny, nx, nz, nt = 40, 44, 48, 40
    vel = smooth_field((ny, nx, nz), 1800.0, 2600.0, 3)
    eps = smooth_field((ny, nx, nz), 0.0, 0.2, 4) if model == VTI else None
    dl = smooth_field((ny, nx, nz), 0.0, 0.1, 5) if model == VTI else None
    rcv = np.array([[x, y, 12.0] for y in range(3, ny - 3, 4) for x in range(3, nx - 3, 4)], np.float32)
    k1, k2 = (0.3, 0.3) if engine == ENGINE_FD else (0.40, 0.48)
    cfg = shot_config(nx, ny, nz, 10.0, 1e-3, nt, engine, order, model, nbnd=8, vmax=2600.0,
                      src_xyz=(21.0, 19.0, 14.0), f_peak=20.0, spread=5, khpass=k1, khcut=k2, nrcv=rcv.shape[0],
                      snap_every=2, snap_format=SNAP_FP32)
    recs = []
    for tuned in (False, True):
        p = Propagator(cfg)
        p.set_model(vel, eps, dl)
        p.set_receivers(rcv)
        if tuned:
            tuner = Autotuner(steps=2)
            best = tuner.tune(p)
            key = next(iter(tuner.cache))
            best2, table = tuner.cache[key]
            print(f"[autotune] {case}: best {best} of {len(table)} candidates", flush=True)
            assert best is not None and best == best2 and best in table
            assert table[best] == min(table.values())
            if engine == ENGINE_CFD:
                assert set(table) == {(bs, ov, 0) for bs in (64, 128, 256) for ov in (0, 1)}
            assert tuner.tune(p) == best            # cached
        recs.append(p.model())
        p.close()
    if engine == ENGINE_CFD:
        assert np.array_equal(recs[0].view(np.uint32), recs[1].view(np.uint32))
    else:
        from tests.helpers import rel_l2
        assert rel_l2(recs[1], recs[0]) < 1e-6
