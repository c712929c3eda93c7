"""Many shots on one GPU: side-by-side shots, tuned stencil tiles, FP64 multi-shot stack.

The reference migrates one shot per process and stacks images on the output
grid (bgprtm/ImagingCondition.cpp:874-945).  Here a batch of shots with their
own aperture sub-grids runs side by side on one GPU: every shot is a
Propagator on its own HIP stream, so the kernels of different shots overlap
(small apertures and the short injection / imaging kernels no longer leave
the GPU idle), and each finished shot image is added in FP64 into the stack
on the full output grid.  A batch holds as many shots as fit the snapshot
budget in HBM.  With ``graph=True`` each shot's time loop is a sequence of
hipGraph replays (one snapshot-interval group per replay, the time index on
a device counter); otherwise the host issues the shots' steps in chunks of
the snapshot cadence, round-robin.  Stencil tile shapes are chosen per
aperture shape by ``Autotuner`` (timed candidates, cached).
"""
import ctypes as C
import time

import numpy as np

from ._lib import check, lib, ENGINE_FD, ISO, SNAP_FP32, SNAP_LEARNED, SNAP_CONVAE
from .propagator import Propagator, shot_config


class Autotuner:
    """Picks (tz, tx, ychunk) for a propagator's stencil step by timing candidates.

    Candidates are z x x tiles of at most 512 threads and y-chunk lengths; each is timed over a few
    forward steps with the library's kernel timers, and the fastest is kept per (grid, engine, order,
    model).  For the compact engine a candidate is (lanes per line block, concurrent passes 0/1, 0).
    The propagator's wavefields are reset afterwards."""

    TILES = ((64, 8), (64, 4), (32, 8), (128, 4))
    CFD_BLOCKS = (64, 128, 256)  # lanes per block of the compact engine's x/y line passes
    CFD_TIMERS = (b"cfd_step",)  # the step's span: its concurrent passes count once
    TILES_FIRST_DERIV = TILES + ((16, 16),)  # the gradient/divergence kernels also take 16 x 16
    YCHUNKS = (0, 16, 32, 64)   # 0: the engine's default chunking

    def __init__(self, steps=4):
        self.steps = steps
        self.cache = {}

    @staticmethod
    def _kernel(p):
        k = C.c_void_p()
        check(lib().klrtm_propagator_kernel(p.h, C.byref(k)))
        return k.value

    def tune(self, p):
        cfg = p.cfg
        key = (cfg.nx, cfg.ny, cfg.nz, cfg.engine, cfg.order, cfg.model)
        k = self._kernel(p)
        if key not in self.cache:
            self.cache[key] = self._search(p, k)
        best = self.cache[key][0]
        if best is not None and cfg.engine != ENGINE_FD:
            check(lib().klrtm_kernel_set_tile(k, best[0], 0, 0))
            check(lib().klrtm_kernel_set_overlap(k, best[1]))
        elif best is not None:
            check(lib().klrtm_kernel_set_tile(k, best[0], best[1], best[2] if best[2] > 0 else -1))
        return best

    def _time(self, p, names):
        """Mean summed kernel time (ms) of one forward step over `self.steps` steps: the named kernels that ran
        (None if none did)."""
        L = lib()
        p.forward_steps(0, 1, False)   # warm
        L.klrtm_synchronize()
        L.klrtm_timer_reset()
        L.klrtm_timer_enable(1)
        p.forward_steps(1, self.steps, False)
        L.klrtm_synchronize()
        L.klrtm_timer_enable(0)
        ms, seen = 0.0, False
        for name in names:
            tot, cnt = C.c_double(), C.c_long()
            L.klrtm_timer_read(name, C.byref(tot), C.byref(cnt))
            if cnt.value:
                ms, seen = ms + tot.value / cnt.value, True
        return ms if seen else None

    def _search(self, p, k):
        L = lib()
        table = {}
        p.reset()
        if p.cfg.engine != ENGINE_FD:   # compact engine: lanes per block of the line passes, concurrent passes
            for bs in self.CFD_BLOCKS:
                for ov in (1, 0):
                    check(L.klrtm_kernel_set_tile(k, bs, 0, 0))
                    check(L.klrtm_kernel_set_overlap(k, ov))
                    ms = self._time(p, self.CFD_TIMERS)
                    if ms is not None:
                        table[(bs, ov, 0)] = ms
        else:
            names = (b"iso_step",) if p.cfg.model == ISO else (b"fd_grad", b"fd_div_step")
            R = p.cfg.order // 2
            for tz, tx in (self.TILES if p.cfg.model == ISO else self.TILES_FIRST_DERIV):
                if tz < R or tx < R or tz * tx > 512:
                    continue
                for yc in self.YCHUNKS:
                    check(L.klrtm_kernel_set_tile(k, tz, tx, yc if yc > 0 else -1))
                    ms = self._time(p, names)
                    if ms is not None:
                        table[(tz, tx, yc)] = ms
        L.klrtm_timer_reset()
        p.reset()
        best = min(table, key=table.get) if table else None
        return best, table


class Shot:
    """One shot of a batch: aperture origin (x0, y0) and size (nx, ny) on the output grid, local
    source and receiver coordinates."""

    def __init__(self, x0, y0, nx, ny, src_xyz, rcv_xyz):
        self.x0, self.y0, self.nx, self.ny = int(x0), int(y0), int(nx), int(ny)
        self.src_xyz = tuple(float(v) for v in src_xyz)
        self.rcv_xyz = np.ascontiguousarray(rcv_xyz, np.float32).reshape(-1, 3)


class BatchRTM:
    """RTM of many shots into one FP64 stack on an output grid of (nyF, nxF, nz).

    velocity(x0, y0, nx, ny) returns the shot's (ny, nx, nz) velocity sub-volume; common holds the
    shot_config keywords shared by all shots (d, dt, nt, engine, order, model, nbnd, ...)."""

    def __init__(self, out_shape, velocity, common, batch=4, snap_format=SNAP_FP32, lcodec=None, tuner=None,
                 graph=False, snap_budget_gb=200.0, convae=None, cae_floor=0.05):
        from ._lib import DeviceArray
        self.nyF, self.nxF, self.nz = out_shape
        self.velocity = velocity
        self.common = dict(common)
        self.batch = max(1, int(batch))
        self.snap_format = snap_format
        self.lcodec = lcodec
        self.convae, self.cae_floor = convae, cae_floor
        self.tuner = tuner
        self.graph = bool(graph)
        self.snap_budget = snap_budget_gb * 1e9
        self.stack = DeviceArray(out_shape, np.float64)
        check(lib().klrtm_synchronize())
        self.streams = []
        for _ in range(self.batch):
            s = C.c_void_p()
            check(lib().klrtm_stream_create(C.byref(s)))
            self.streams.append(s.value)
        self.times = dict(migrate_s=0.0, shots=0, steps=0)

    def _make(self, shot):
        c = dict(self.common)
        nt = c.pop("nt")
        vel = self.velocity(shot.x0, shot.y0, shot.nx, shot.ny)
        c.setdefault("vmax", float(vel.max()))
        cfg = shot_config(shot.nx, shot.ny, self.nz, c.pop("d"), c.pop("dt"), nt, nrcv=shot.rcv_xyz.shape[0],
                          src_xyz=shot.src_xyz, snap_format=self.snap_format, ic_fp32_only=1, **c)
        p = Propagator(cfg)
        if self.snap_format == SNAP_LEARNED:
            check(lib().klrtm_propagator_set_lcodec(p.h, self.lcodec.h))
        elif self.snap_format == SNAP_CONVAE:
            check(lib().klrtm_propagator_set_convae(p.h, self.convae.h, C.c_float(self.cae_floor)))
        p.set_model(vel)
        p.set_receivers(shot.rcv_xyz)
        return p

    def snapshot_bytes(self, shot):
        """HBM bytes of one shot's snapshots (every snap_every-th of nt steps, slots padded to 256 B)."""
        nt, k = self.common["nt"], max(1, self.common.get("snap_every", 1))
        shape = (shot.ny, shot.nx, self.nz)
        if self.snap_format == SNAP_LEARNED:
            per = self.lcodec.code_bytes(shape)
        elif self.snap_format == SNAP_CONVAE:
            per = int(lib().klrtm_convae_code_bytes(shot.nx, shot.ny, self.nz))
        elif self.snap_format == SNAP_FP32:
            per = 4 * shot.nx * shot.ny * self.nz
        else:
            per = 2 * shot.nx * shot.ny * self.nz
        return ((nt + k - 1) // k) * ((per + 255) // 256 * 256)

    def groups(self, shots):
        """Consecutive shots, at most `batch` per group and within the snapshot budget (a shot over the
        budget on its own still forms a group)."""
        out, cur, used = [], [], 0
        for sh in shots:
            b = self.snapshot_bytes(sh)
            if cur and (len(cur) == self.batch or used + b > self.snap_budget):
                out.append(cur)
                cur, used = [], 0
            cur.append(sh)
            used += b
        if cur:
            out.append(cur)
        return out

    def run(self, shots, records=None):
        """Migrate the shots; records[i] is shot i's [nrcv][nt] record (None: modelled here, untimed).
        Returns the per-shot records used."""
        L = lib()
        used = []
        self.times.setdefault("groups", [])
        g0 = 0
        for group in self.groups(shots):
            self.times["groups"].append(len(group))
            props = []
            for j, sh in enumerate(group):
                L.klrtm_set_stream(self.streams[j])
                p = self._make(sh)
                if self.tuner is not None:
                    self.tuner.tune(p)
                rec = records[g0 + j] if records is not None else p.model()
                p.set_record(rec)
                used.append(rec)
                p.forward_steps(0, 1, True)   # allocates the snapshot store outside the timed region
                if self.graph:
                    p.set_graph(True)
                props.append(p)
            for j in range(len(group)):
                L.klrtm_stream_synchronize(self.streams[j])
            t0 = time.perf_counter()
            nt = self.common["nt"]
            k = max(1, self.common.get("snap_every", 1))
            for j, p in enumerate(props):
                L.klrtm_set_stream(self.streams[j])
                p.reset()
            if self.graph:  # each shot's whole forward run, then its backward run, as graph replays
                for j, p in enumerate(props):
                    L.klrtm_set_stream(self.streams[j])
                    p.forward_steps(0, nt, True)
                for j, p in enumerate(props):
                    L.klrtm_set_stream(self.streams[j])
                    p.reset()
                    check(L.klrtm_propagator_clear_images(p.h))
                    p.backward_steps(nt - 1, nt, True)
            else:
                for it0 in range(0, nt, k):
                    for j, p in enumerate(props):
                        L.klrtm_set_stream(self.streams[j])
                        p.forward_steps(it0, min(k, nt - it0), True)
                for j, p in enumerate(props):
                    L.klrtm_set_stream(self.streams[j])
                    p.reset()
                    check(L.klrtm_propagator_clear_images(p.h))
                it = nt - 1
                while it >= 0:
                    n = it % k + 1  # down to the next imaged step
                    for j, p in enumerate(props):
                        L.klrtm_set_stream(self.streams[j])
                        p.backward_steps(it, n, True)
                    it -= n
            # apertures overlap: the stack is updated shot after shot on one stream, in shot order
            for j in range(len(group)):
                check(L.klrtm_stream_synchronize(self.streams[j]))
            L.klrtm_set_stream(self.streams[0])
            for p, sh in zip(props, group):
                img = C.c_void_p()
                check(L.klrtm_propagator_images(p.h, C.byref(img), None, None, None))
                check(L.klrtm_stack_image(img.value, 0, sh.nx, sh.ny, self.nz, self.stack.ptr, self.nxF, self.nyF,
                                          self.nz, sh.x0, sh.y0, 0))
            check(L.klrtm_stream_synchronize(self.streams[0]))
            self.times["migrate_s"] += time.perf_counter() - t0
            self.times["shots"] += len(group)
            self.times["steps"] += len(group) * nt
            for p in props:
                p.close()
            g0 += len(group)
        L.klrtm_set_stream(None)
        return used

    def shots_per_hour(self, nt_full=None):
        """Shots per hour of migration (forward + backward + imaging), scaled to nt_full steps per shot
        when the measured shots ran fewer."""
        if not self.times["shots"]:
            return 0.0
        per_shot = self.times["migrate_s"] / self.times["shots"]
        if nt_full:
            per_shot *= nt_full / self.common["nt"]
        return 3600.0 / per_shot

    def close(self):
        for s in self.streams:
            lib().klrtm_stream_destroy(s)
        self.streams = []
