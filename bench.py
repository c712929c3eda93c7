#!/usr/bin/env python3
"""bench.py -- RTM propagation-and-imaging throughput on one MI355X.

One benchmark step is one RTM time step of a shot: a forward step of the
source wavefield (kernel + leapfrog + source injection + snapshot store on
its cadence) and a backward step of the receiver wavefield (kernel + leapfrog
+ injection of the full surface receiver grid + zero-lag imaging condition on
the snapshot cadence).  Inputs are resident in HBM when the timed region
starts.  `value` counts both propagations: 2 * nx*ny*nz cell updates per step.

Default workload: VTI pseudo-acoustic RTM with the
10th-order compact FD engine on a 400^3 grid, dx = 12.5 m, 32-cell PML,
20 Hz Ricker, 400x400 surface receivers, FP32 snapshots every 4 steps.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s + RTM shots/hour, 1 GPU, learned/batched vs plain at equal NRMSE"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: grid (ny, nx, nz), spacing, engine, order, model, PML, Ricker fpeak, snapshot cadence, record length
    "vti-cfd10-400": dict(shape=(400, 400, 400), d=12.5, engine="CFD", order=10, model="VTI", nbnd=32, fpeak=20.0,
                          snap_every=4, record_s=2.0, seed=202),
    "iso-fd8-400": dict(shape=(400, 400, 400), d=12.5, engine="FD", order=8, model="ISO", nbnd=32, fpeak=15.0,
                        snap_every=4, record_s=2.0, seed=505),
    "iso-fd8-128": dict(shape=(128, 128, 128), d=10.0, engine="FD", order=8, model="ISO", nbnd=20, fpeak=15.0,
                        snap_every=1, record_s=0.6, seed=101, dt=1e-3),
    "tti-fd16-600": dict(shape=(600, 600, 400), d=12.5, engine="FD", order=16, model="TTI", nbnd=40, fpeak=15.0,
                         snap_every=3, record_s=2.325, seed=303),
}
# algorithmic HBM bytes per cell update of the step (state + model arrays read once, new field written once)
VOL_BYTES = {"ISO": 16, "VTI": 24, "TTI": 32}


def build_case(name):
    from konglong_amd import models
    from konglong_amd._lib import ENGINE_FD, ENGINE_CFD, ISO, VTI, TTI
    w = dict(WORKLOADS[name])
    ny, nx, nz = w["shape"]
    eng = ENGINE_CFD if w["engine"] == "CFD" else ENGINE_FD
    mdl = {"ISO": ISO, "VTI": VTI, "TTI": TTI}[w["model"]]
    eps = dl = None
    if mdl == ISO and nz == 128:
        vel = models.layered_model(nz, w["d"], w["seed"])
    elif mdl == ISO:
        vel = models.smooth_vti_model(w["shape"], 20.0, w["seed"])[0]
    else:
        vel, eps, dl = models.smooth_vti_model(w["shape"], 20.0, w["seed"])
    if mdl == TTI:  # tilt 0-45 deg, azimuth 0-360 deg, smoothed like the other fields
        w["pj"] = models.tilt_axes(w["shape"], 20.0, w["seed"] + 10)
    dt = w.get("dt") or models.cfl_dt(vel, w["d"], eng, w["order"], mdl, eps, dl, *w.get("pj", (None, None)))
    return w, eng, mdl, vel, eps, dl, dt


def parse_traffic(name):
    """HBM bytes per step from the committed rocprofv3 summary of this workload (or None)."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except ValueError:
                pass
    if "fetch_bytes_per_step" in vals and "write_bytes_per_step" in vals:
        return vals["fetch_bytes_per_step"] + vals["write_bytes_per_step"]
    return None


def _cpu_info():
    """Host CPU model, logical CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def _timing_oracle():
    """The oracle built like the reference (-O3 -ffast-math, bgprtm make.mk:63-66): for this host's own ISA
    (-march=native, compiled here in ~10 s) when a compiler is present, else the prebuilt x86-64-v3 build."""
    import subprocess
    odir = os.path.join(ROOT, "oracle")
    try:
        subprocess.run(["make", "-s", "-C", odir, "liboracle_native.so"], check=True, timeout=120,
                       capture_output=True)
        return os.path.join(odir, "liboracle_native.so"), "-O3 -ffast-math -march=native -fopenmp"
    except (OSError, subprocess.SubprocessError):
        return os.path.join(odir, "liboracle_fast.so"), "-O3 -ffast-math -march=x86-64-v3 -fopenmp"


def cpu_baseline(vel, eps, dl, w, eng, mdl, dt, budget_s, nt_shot):
    """The FP32 CPU oracle (reference-structured, OpenMP over every CPU this process may use) timed on this
    host for a bounded sample of time steps; the per-step rate is the reported value."""
    import oracle
    ny, nx, nz = w["shape"]
    model, ncpu, usable = _cpu_info()
    threads = usable
    path, flags = _timing_oracle()
    oracle.lib(path)
    m = (vel * vel * (np.float32(dt) * np.float32(dt))).astype(np.float32)
    o = oracle.Engine(nx, ny, nz, w["d"], w["d"], w["d"], eng, w["order"], np.float32, threads=threads)
    o.set_model(m, eps, dl, *w.get("pj", (None, None)))
    o.set_boundary([w["nbnd"]] * 6, dt, float(vel.max()), dim_pml=3)
    rng = np.random.default_rng(1)
    w1 = (1e-3 * rng.standard_normal(w["shape"])).astype(np.float32)
    w0, wr = w1.copy(), np.zeros_like(w1)
    t = time.perf_counter()
    o.step(mdl, w1, w0, wr)
    one = time.perf_counter() - t
    n = max(1, min(40, int(budget_s / max(one, 1e-3))))
    t = time.perf_counter()
    for _ in range(n):
        o.step(mdl, w1, w0, wr)
        w0, w1 = w1, w0
    el = time.perf_counter() - t
    o.close()
    rate = nx * ny * nz * n / el / 1e9
    return {"value": round(rate, 5), "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} {w['model']}-{w['engine']}{w['order']} time steps (kernel + leapfrog, PML) on the "
                      f"{ny}x{nx}x{nz} model after one untimed step, FP32, {threads} OpenMP threads; the shot rate "
                      f"extrapolates this step rate to the 2 x nt_shot steps of one RTM shot",
            "cpu_model": model, "nproc": ncpu, "usable_cpus": usable, "build_flags": flags,
            "shots_per_hour_extrapolated": round(3600.0 / (2 * nt_shot * el / n), 4)}


def adjacent_physics(steps=6):
    """Kernel timings of the paths either side of the step on the 400^3 grid (dx 12.5 m, PML 32): the Q
    leapfrog, density on the first-derivative form, ghost-mirrored injection, the snapshot cache in host
    memory and the FFT time shift.  Times are HIP-event means per launch (ms) or wall times (s)."""
    import ctypes as C
    from konglong_amd import models
    from konglong_amd._lib import ENGINE_FD, ISO, VTI, SNAP_FP32, DeviceArray, lib, check, fptr
    from konglong_amd.propagator import Propagator, shot_config
    from konglong_amd.receiver import Receiver
    L = lib()
    shape = (400, 400, 400)
    ny, nx, nz = shape
    n = nx * ny * nz
    vel, eps, dl = models.smooth_vti_model(shape, 20.0, 606)
    dt = models.cfl_dt(vel, 12.5, ENGINE_FD, 8, VTI, eps, dl)
    rcv = models.surface_receivers(nx, ny, 34)

    def timed(p, names, nsteps=steps):
        p.reset()
        p.forward_steps(0, 1, False)
        L.klrtm_synchronize()
        check(L.klrtm_timer_reset())
        check(L.klrtm_timer_enable(1))
        p.forward_steps(1, nsteps, False)
        L.klrtm_synchronize()
        check(L.klrtm_timer_enable(0))
        out = {}
        for nm in names:
            t, c = C.c_double(), C.c_long()
            L.klrtm_timer_read(nm.encode(), C.byref(t), C.byref(c))
            if c.value:
                out[nm] = round(t.value / c.value, 4)
        return out

    def shot(model, nt=2 + steps, **kw):
        cfg = shot_config(nx, ny, nz, 12.5, dt, nt, ENGINE_FD, 8, model, nbnd=32, vmax=float(vel.max()),
                          src_xyz=(199.5, 199.5, 34.0), f_peak=15.0, spread=10, nrcv=rcv.shape[0], snap_every=4,
                          snap_format=SNAP_FP32, **kw)
        p = Propagator(cfg)
        p.set_receivers(rcv)
        return p

    res = {"grid": list(shape)}
    # Q (order 3) on the ISO FD8 step: kernel + Q leapfrog against the fused step
    p = shot(ISO)
    p.set_model(vel)
    base = timed(p, ["iso_step"])
    p.set_q(np.full(shape, 0.02, np.float32), np.array([0.04, 0.3, 0.2, 0.1, 0.9, 0.7, 0.5], np.float32))
    p.set_model(vel)
    q = timed(p, ["iso_pb", "update2nd_q"])
    res["q_order3"] = {"iso_step_ms": base.get("iso_step"), **{k + "_ms": v for k, v in q.items()},
                       "update2nd_q_GBps": round(n * 52 / (q["update2nd_q"] * 1e-3) / 1e9, 1)}
    p.close()
    # density on the VTI FD8 step (gradient + divergence kernels), against no density
    p = shot(VTI)
    p.set_model(vel, eps, dl)
    a = timed(p, ["fd_grad", "fd_div_step"])
    p.set_rho(models.smooth_vti_model(shape, 20.0, 607)[0] / np.float32(2000.0))
    b = timed(p, ["fd_grad", "fd_div_step"])
    res["density_vti_fd8"] = {"without_ms": round(sum(a.values()), 4), "with_ms": round(sum(b.values()), 4)}
    p.close()
    # ghost-mirrored injection of the 160 000-receiver surface (spread 10)
    r = Receiver(nx, ny, nz, rcv, 10, 0.3, 0.3, 4, 1.0 / 12.5 ** 3)
    r.set_traces(np.random.default_rng(1).standard_normal((rcv.shape[0], 4)).astype(np.float32))
    w, dv = DeviceArray(shape), DeviceArray.from_host(vel)
    gh = {}
    for mirror in (-1, 24):
        r.apply(w, 0, dv, 1.0, mirror)
        L.klrtm_synchronize()
        check(L.klrtm_timer_reset())
        check(L.klrtm_timer_enable(1))
        for it in range(4):
            r.apply(w, it, dv, 1.0, mirror)
        L.klrtm_synchronize()
        check(L.klrtm_timer_enable(0))
        t, c = C.c_double(), C.c_long()
        L.klrtm_timer_read(b"inject", C.byref(t), C.byref(c))
        gh["mirror" if mirror >= 0 else "plain"] = round(t.value / c.value, 4)
    res["injection_ms"] = gh
    del w, dv, r
    # snapshot cache in host memory: a 40-step RTM with 10 FP32 snapshots (2.56 GB) against HBM
    times = {}
    for cached in (False, True):
        p = shot(ISO, nt=40)
        p.set_model(vel)
        if cached:
            p.set_snapshot_cache(None, mem_limit_gb=8.0)
        rec = np.zeros((rcv.shape[0], 40), np.float32)
        p.rtm(rec, want=("image",))  # warm (allocations, pinned staging)
        t0 = time.perf_counter()
        p.rtm(rec, want=("image",))
        times["host_cache" if cached else "hbm"] = round(time.perf_counter() - t0, 3)
        p.close()
    res["rtm_40_steps_s"] = times
    # FFT time shift: 1000 traces of 3000 samples
    tr = DeviceArray.from_host(np.random.default_rng(2).standard_normal(3000).astype(np.float32))
    out = DeviceArray((1000, 3000))
    sh = np.linspace(-50, 50, 1000).astype(np.float32)
    nfft = int(L.klrtm_fft_size(3000))
    check(L.klrtm_timeshift(tr.ptr, 3000, nfft, 1000, fptr(sh), out.ptr))
    t0 = time.perf_counter()
    check(L.klrtm_timeshift(tr.ptr, 3000, nfft, 1000, fptr(sh), out.ptr))
    res["timeshift_1000x3000_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    return res


def batched_vs_plain(nshots, nt, batch, seed=505):
    """Batched multi-shot ISO RTM against the plain per-shot FP32 path on the same shots.

    Workload: the first `nshots` of 64 seeded apertures (200-450 cells square, multiples of 8, x 400
    deep) on a 1000 x 1000 x 400 layered model with a 4500 m/s salt body, dx 12.5 m, FD8, PML 32,
    15 Hz Ricker, receivers every 4th surface cell, snapshots every 4 steps kept in HBM, nt steps per
    shot (shots/hour is scaled to 2500 steps when nt is shorter).  Plain: one shot at a time, FP32
    snapshots, default tiles, launches issued one by one.  Batched: up to `batch` shots side by side
    on their own streams (as many as fit a 200 GB snapshot budget), each shot's time loop replayed as
    captured hipGraphs, autotuned tiles; FP32 snapshots, then learned snapshots (trained beforehand,
    untimed, on two seeded 256^3 random models).  All stack FP32 shot images into an FP64 image; the
    batched stacks' NRMSE is taken against the plain one (the learned path's NRMSE against an FP64
    oracle is measured by the batched-shot GPU test on small apertures)."""
    from konglong_amd import models, learned
    from konglong_amd.batch import BatchRTM, Shot, Autotuner
    from konglong_amd._lib import ENGINE_FD, ISO, SNAP_FP32, SNAP_LEARNED
    nxF = nyF = 1000
    nz, d, nb, fp, k = 400, 12.5, 32, 15.0, 4
    model = models.SaltLayeredModel(nxF, nyF, nz, d, seed)
    dt = 2.0 / models.fd_k1max(8) / (4500.0 * np.sqrt(3.0) / d) / 1.01   # bgprtm/ModelPrepare.cpp:701-718 rule
    rng = np.random.default_rng(seed)
    shots = []
    for i in range(64):
        a = (int(rng.uniform(200, 451)) // 8 * 8, int(rng.uniform(200, 451)) // 8 * 8)
        x0, y0 = int(rng.uniform(0, nxF - a[0])), int(rng.uniform(0, nyF - a[1]))
        rx = models.surface_receivers(a[0], a[1], nb + 2, step=4)
        keep = (rx[:, 0] >= nb) & (rx[:, 0] < a[0] - nb) & (rx[:, 1] >= nb) & (rx[:, 1] < a[1] - nb)
        shots.append(Shot(x0, y0, a[0], a[1], ((a[0] - 1) / 2, (a[1] - 1) / 2, nb + 2.0), rx[keep]))
    shots = shots[:nshots]
    cache = {}

    def vel(x0, y0, nx, ny):
        key = (x0, y0, nx, ny)
        if key not in cache:
            cache[key] = model(x0, y0, nx, ny)
        return cache[key]
    common = dict(d=d, dt=float(dt), nt=nt, engine=ENGINE_FD, order=8, model=ISO, nbnd=nb, f_peak=fp, spread=10,
                  snap_every=k, tmute_frac=0.1, vmax=4500.0)
    t0 = time.perf_counter()
    codec, info = learned.train((256, 256, 256), [4001, 4002], d, float(dt), 300, k, nl=32, nbnd=nb, f_peak=fp,
                                vmin=1500.0, vmax=4500.0, sigma=20.0)
    t_train = time.perf_counter() - t0
    plain = BatchRTM((nyF, nxF, nz), vel, common, batch=1, snap_format=SNAP_FP32)
    records = plain.run(shots)
    print(f"[bench] batched leg: plain {nshots} shots x {nt} steps done", file=sys.stderr, flush=True)
    tuner = Autotuner()
    fast32 = BatchRTM((nyF, nxF, nz), vel, common, batch=batch, snap_format=SNAP_FP32, tuner=tuner, graph=True)
    fast32.run(shots, records)
    print("[bench] batched leg: batched FP32 done", file=sys.stderr, flush=True)
    fast = BatchRTM((nyF, nxF, nz), vel, common, batch=batch, snap_format=SNAP_LEARNED, lcodec=codec, tuner=tuner,
                    graph=True)
    fast.run(shots, records)
    a64 = plain.stack.download()
    na = max(np.linalg.norm(a64), 1e-300)
    err = float(np.linalg.norm(fast.stack.download() - a64) / na)
    err32 = float(np.linalg.norm(fast32.stack.download() - a64) / na)
    sph = [b.shots_per_hour(2500) for b in (plain, fast32, fast)]
    ratio = codec.code_bytes((nz, nz, nz)) / (nz ** 3 * 4.0)
    for b in (plain, fast32, fast, codec):
        b.close()
    return {"workload": "batched ISO FD8 RTM, 1000x1000x400 salt model, apertures 200-450 x 400",
            "shots": nshots, "steps_per_shot_measured": nt, "steps_per_shot_scaled_to": 2500, "batch": batch,
            "apertures": [[sh.nx, sh.ny] for sh in shots], "snapshot_gb_per_shot_fp32":
                [round(plain.snapshot_bytes(sh) / 1e9, 1) for sh in shots],
            "groups_fp32": fast32.times["groups"], "groups_learned": fast.times["groups"],
            "time_loop": "plain: launches issued one by one; batched: hipGraph replays of snapshot-interval groups",
            "plain_shots_per_hour": round(sph[0], 2),
            "batched_fp32_shots_per_hour": round(sph[1], 2), "batched_fp32_speedup": round(sph[1] / sph[0], 3),
            "batched_fp32_stack_nrmse_vs_plain": err32,
            "batched_learned_shots_per_hour": round(sph[2], 2), "batched_learned_speedup": round(sph[2] / sph[0], 3),
            "batched_learned_stack_nrmse_vs_plain": err, "learned_bytes_vs_fp32": round(ratio, 4),
            "codec_training_s": round(t_train, 2), "codec_training": info,
            "tiles": {f"{k[0]}x{k[1]}": list(v[0]) if v[0] else None for k, v in tuner.cache.items()}}


def convae_vs_plain(train_models=8, train_steps=20000, batch=64, floor_frac=0.05, seed=606):
    """Convolutional-autoencoder snapshots inside a whole 400^3 ISO FD8 RTM shot, against FP32, the
    reference's COMPRESS_16 (the WaveFieldCompress fallback) and the linear learned codec.

    Training (untimed in the shot rates): Adam (lr 1e-3, batches of `batch` 32^3 blocks, `train_steps`
    steps) on 4 forward snapshots of one shot on each of `train_models` seeded random velocity models at
    256^3 (dx 12.5 m, 1500-5000 m/s).  Then one whole shot of the iso-fd8-400 workload (2 s record,
    receivers on every surface cell, snapshots every 4 steps, all kept in HBM) per snapshot format:
    shots/hour, snapshot bytes against FP32, image NRMSE against the FP32-snapshot image (the NRMSE
    against an FP64 oracle of the same kind of codec is measured by the autoencoder GPU test at 64^3)."""
    from konglong_amd import convae, learned, models
    from konglong_amd._lib import ENGINE_FD, ISO, SNAP_FP32, SNAP_C16, SNAP_LEARNED, SNAP_CONVAE, lib, check
    from konglong_amd.propagator import Propagator, shot_config
    w, eng, mdl, vel, eps, dl, dt = build_case("iso-fd8-400")
    ny, nx, nz = w["shape"]
    d, nb, fp, k = w["d"], w["nbnd"], w["fpeak"], w["snap_every"]
    nt = int(round(w["record_s"] / dt))
    t0 = time.perf_counter()
    bank = convae.snapshot_bank((256, 256, 256), [seed + i for i in range(train_models)], d, dt, 1000, per_model=4,
                                nbnd=nb, f_peak=fp, vmin=1500.0, vmax=5000.0, sigma=20.0)
    t_bank = time.perf_counter() - t0
    net = convae.ConvAutoencoder(seed)
    t0 = time.perf_counter()
    l0, l1 = net.fit(bank, train_steps, batch, 1e-3, floor_frac, seed)
    _lib_sync()
    t_train = time.perf_counter() - t0
    bank.free()
    print(f"[bench] autoencoder trained: loss {l0:.4f} -> {l1:.4f} in {t_train:.1f} s", file=sys.stderr, flush=True)
    lin, _ = learned.train((256, 256, 256), [seed + 100, seed + 101], d, dt, 300, k, nl=32, nbnd=nb, f_peak=fp,
                           vmin=1500.0, vmax=5000.0, sigma=20.0)
    # equal bytes: snapshot reconstruction error on held-out models, the autoencoder (4100 B per 32^3
    # block) against the linear code without its int8 residual (nl = 32 BF16 values per 8^3 block:
    # 4096 B per 32^3 of latents, residual scales zeroed) and with it
    held = convae.snapshot_bank((256, 256, 256), [seed + 200, seed + 201], d, dt, 1000, per_model=2, nbnd=nb,
                                f_peak=fp, vmin=1500.0, vmax=5000.0, sigma=20.0)
    shp = (256, 256, 256)
    nvol = 256 ** 3
    from konglong_amd._lib import DeviceArray
    cae_code = DeviceArray((convae.ConvAutoencoder.code_bytes(shp),), np.uint8)
    lin_code = lin.new_code(shp)
    snap, dec = DeviceArray(shp), DeviceArray(shp)
    nblk8 = (256 // 8) ** 3
    zsz = (nblk8 * 32 * 2 + 255) // 256 * 256
    errs = {"autoencoder": [], "linear_no_residual": [], "linear": []}
    for i in range(held.shape[0]):
        check(lib().klrtm_memcpy_d2d(snap.ptr, held.ptr + i * nvol * 4, nvol * 4))
        ref_v = snap.download().astype(np.float64)
        nref = np.linalg.norm(ref_v)
        net.encode(snap, cae_code, floor_frac)
        net.decode(cae_code, dec)
        errs["autoencoder"].append(float(np.linalg.norm(dec.download() - ref_v) / nref))
        lin.encode(snap, lin_code)
        lin.decode(lin_code, dec)
        errs["linear"].append(float(np.linalg.norm(dec.download() - ref_v) / nref))
        check(lib().klrtm_memset(lin_code.ptr + zsz, 0, nblk8 * 4))  # residual scales -> 0
        lin.decode(lin_code, dec)
        errs["linear_no_residual"].append(float(np.linalg.norm(dec.download() - ref_v) / nref))
    held.free()
    equal_bytes = {"snapshots": len(errs["linear"]),
                   "bytes_vs_fp32": {"autoencoder": round(convae.ConvAutoencoder.code_bytes(shp) / (4.0 * nvol), 4),
                                     "linear_no_residual": round(nblk8 * 32 * 2 / (4.0 * nvol), 4),
                                     "linear": round(lin.code_bytes(shp) / (4.0 * nvol), 4)},
                   "mean_rel_l2": {k_: round(float(np.mean(v)), 5) for k_, v in errs.items()}}
    rcv = models.surface_receivers(nx, ny, nb + 2)
    common = dict(nbnd=nb, vmax=float(vel.max()), src_xyz=((nx - 1) / 2, (ny - 1) / 2, nb + 2.0), f_peak=fp,
                  spread=10, nrcv=rcv.shape[0], snap_every=k, tmute_frac=0.1, ic_fp32_only=1)
    rec = None
    res = {}
    imgs = {}
    for name, fmt in (("fp32", SNAP_FP32), ("compress16", SNAP_C16), ("linear", SNAP_LEARNED),
                      ("autoencoder", SNAP_CONVAE)):
        p = Propagator(shot_config(nx, ny, nz, d, dt, nt, ENGINE_FD, 8, ISO, snap_format=fmt, **common))
        if fmt == SNAP_LEARNED:
            check(lib().klrtm_propagator_set_lcodec(p.h, lin.h))
        elif fmt == SNAP_CONVAE:
            check(lib().klrtm_propagator_set_convae(p.h, net.h, floor_frac))
        p.set_model(vel)
        p.set_receivers(rcv)
        if rec is None:
            rec = p.model()
        p.forward_steps(0, 1, True)  # allocates the snapshot store outside the timed shot
        t0 = time.perf_counter()
        imgs[name] = p.rtm(rec, want=("image",))["image"]
        t_shot = time.perf_counter() - t0
        code = {SNAP_FP32: 4 * nx * ny * nz, SNAP_C16: 2 * nx * ny * nz + 4 * ny,
                SNAP_LEARNED: lin.code_bytes((ny, nx, nz)),
                SNAP_CONVAE: convae.ConvAutoencoder.code_bytes((ny, nx, nz))}[fmt]
        res[name] = {"shots_per_hour": round(3600.0 / t_shot, 2), "shot_s": round(t_shot, 3),
                     "snapshot_bytes_vs_fp32": round(code / (4.0 * nx * ny * nz), 4)}
        p.close()
        print(f"[bench] convae leg: {name} shot {t_shot:.2f} s", file=sys.stderr, flush=True)
    ref = imgs["fp32"].astype(np.float64)
    for name in res:
        res[name]["image_nrmse_vs_fp32"] = float(np.linalg.norm(imgs[name] - ref) / np.linalg.norm(ref))
    net.close()
    lin.close()
    return {"workload": "iso-fd8-400 whole shot, autoencoder vs FP32 / COMPRESS_16 / linear snapshots",
            "nt": nt, "snap_every": k, "train_models_256": train_models, "train_steps": train_steps, "batch": batch,
            "train_s": round(t_train, 2), "train_ms_per_step": round(t_train / max(train_steps, 1) * 1e3, 3),
            "bank_s": round(t_bank, 2), "loss_first": l0, "loss_last": l1,
            "snapshot_reconstruction_equal_bytes": equal_bytes, "formats": res}


def _lib_sync():
    from konglong_amd import _lib
    _lib.synchronize()


def tti_full_shot(nt=3000, seed=303):
    """One whole TTI FD16 RTM shot on the 600 x 600 x 400 grid (dx 12.5 m, PML 40, 15 Hz, snapshots every
    3 steps), with learned snapshots held in HBM for the whole shot (163 GB for the 1000 snapshots; FP16
    with block scales would need 289 GB, FP32 577 GB).  Returns the shot's timings."""
    from konglong_amd import models, learned
    from konglong_amd._lib import ENGINE_FD, TTI, SNAP_LEARNED, lib, check
    from konglong_amd.propagator import Propagator, shot_config
    shape = (600, 600, 400)
    ny, nx, nz = shape
    d, nb, fp, k = 12.5, 40, 15.0, 3
    t0 = time.perf_counter()
    vel, eps, dl = models.smooth_vti_model(shape, 20.0, seed)
    pjx, pjy = models.tilt_axes(shape, 20.0, seed + 10)
    dt = models.cfl_dt(vel, d, ENGINE_FD, 16, TTI, eps, dl, pjx, pjy)
    t_model = time.perf_counter() - t0
    codec, info = learned.train((256, 256, 256), [4011, 4012], d, dt, 300, k, nl=32, nbnd=nb, f_peak=fp,
                                vmin=1500.0, vmax=5000.0, sigma=20.0)
    rcv = models.surface_receivers(nx, ny, nb + 2, step=4)
    keep = (rcv[:, 0] >= nb) & (rcv[:, 0] < nx - nb) & (rcv[:, 1] >= nb) & (rcv[:, 1] < ny - nb)
    rcv = rcv[keep]
    common = dict(nbnd=nb, vmax=float(vel.max()), src_xyz=((nx - 1) / 2, (ny - 1) / 2, nb + 2.0), f_peak=fp,
                  spread=10, nrcv=rcv.shape[0], snap_every=k, tmute_frac=0.1, ic_fp32_only=1)
    p = Propagator(shot_config(nx, ny, nz, d, dt, nt, ENGINE_FD, 16, TTI, snap_format=SNAP_LEARNED, **common))
    check(lib().klrtm_propagator_set_lcodec(p.h, codec.h))
    p.set_model(vel, eps, dl, pjx, pjy)
    p.set_receivers(rcv)
    t0 = time.perf_counter()
    rec = p.model()
    t_mod = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = p.rtm(rec, want=("image",))
    t_rtm = time.perf_counter() - t0
    tm = p.timings()
    p.close()
    codec.close()
    img = out["image"]
    snap_gb = (nx // 8) * (ny // 8) * (nz // 8) * 580.74 * (nt // k) / 1e9
    return {"metric": METRIC, "value": round(3600.0 / t_rtm, 3), "unit": "shots/hour", "n_gpus": 1,
            "higher_is_better": True, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "tti-fd16-600 whole shot", "grid": [nx, ny, nz], "spacing_m": d, "order": 16,
                       "pml": nb, "dt_s": round(dt, 7), "nt": nt, "snap_every": k, "snap_format": "learned",
                       "receivers": int(rcv.shape[0])},
            "rtm_s": round(t_rtm, 2), "modelling_s": round(t_mod, 2), "forward_ms": round(tm["forward_ms"], 1),
            "backward_ms": round(tm["backward_ms"], 1), "learned_snapshot_store_gb": round(snap_gb, 1),
            "fp16_snapshot_store_gb": round(nx * ny * nz * 2.0 * (1 + 2.0 / 512) * (nt // k) / 1e9, 1),
            "image_finite": bool(np.isfinite(img).all()),
            "image_rms": float(np.sqrt(np.mean(img.astype(np.float64) ** 2))), "model_build_s": round(t_model, 1),
            "codec_training": info}


def measure(workload, K, W, cpu=False, cpu_budget=15.0, overlap=True):
    """One workload: K timed forward+backward RTM steps after W warm-up steps; the bench line's fields."""
    import ctypes as C
    from konglong_amd import _lib
    from konglong_amd import models
    from konglong_amd.propagator import Propagator, shot_config
    from konglong_amd._lib import SNAP_FP32

    L = _lib.lib()
    w, eng, mdl, vel, eps, dl, dt = build_case(workload)
    ny, nx, nz = w["shape"]
    ncell = nx * ny * nz
    nb = w["nbnd"]
    nt_shot = int(round(w["record_s"] / dt))
    nt = W + K + 2
    iz_rcv = nb + 2 if nb > 0 else 2
    rcv = models.surface_receivers(nx, ny, iz_rcv)
    kband = (0.40, 0.48) if eng == _lib.ENGINE_CFD else (0.3, 0.3)
    cfg = shot_config(nx, ny, nz, w["d"], dt, nt, eng, w["order"], mdl, nbnd=nb, vmax=float(vel.max()),
                      src_xyz=((nx - 1) / 2, (ny - 1) / 2, float(iz_rcv)), f_peak=w["fpeak"], spread=10,
                      khpass=kband[0], khcut=kband[1], nrcv=rcv.shape[0], snap_every=w["snap_every"],
                      snap_format=SNAP_FP32, tmute_frac=0.1, fuse_step=1, snap_slots=16)
    p = Propagator(cfg)
    p.set_model(vel, eps, dl, *w.get("pj", (None, None)))
    p.set_receivers(rcv)
    if not overlap:   # compact engine: its line passes one after another
        kh = C.c_void_p()
        _lib.check(L.klrtm_propagator_kernel(p.h, C.byref(kh)))
        _lib.check(L.klrtm_kernel_set_overlap(kh, 0))
    rng = np.random.default_rng(7)
    p.set_record((1e-3 * rng.standard_normal((rcv.shape[0], nt))).astype(np.float32))

    # warm-up: forward then backward
    p.forward_steps(0, W, store_snapshots=True)
    p.backward_steps(W - 1, W, image=True)
    _lib.synchronize()
    # timed region: K forward + K backward steps; HIP events around every launch
    _lib.check(L.klrtm_timer_reset())
    _lib.check(L.klrtm_timer_enable(1))
    _lib.synchronize()
    t0 = time.perf_counter()
    p.forward_steps(W, K, store_snapshots=True)
    p.backward_steps(W + K - 1, K, image=True)
    _lib.synchronize()
    el = time.perf_counter() - t0
    _lib.check(L.klrtm_timer_enable(0))

    names = ["iso_step", "iso_kernel", "update2nd", "fd_grad", "fd_div", "fd_div_step", "cfd_grad_y", "cfd_grad_z",
             "cfd_grad_x", "cfd_div_x", "cfd_div_z", "cfd_div_y", "cfd_step", "inject", "extract", "c16_encode",
             "blockq_encode", "ic_correlate"]
    kern = {}
    tot, cnt = C.c_double(), C.c_long()
    for k in names:
        L.klrtm_timer_read(k.encode(), C.byref(tot), C.byref(cnt))
        if cnt.value:
            kern[k] = {"mean_ms": round(tot.value / cnt.value, 4), "launches": cnt.value}

    ms_step = el / K * 1e3
    value = 2.0 * ncell * K / el / 1e9
    # dominant kernel: the propagation step (one fused launch for FD ISO, the line-solve passes for CFD)
    if eng == _lib.ENGINE_CFD:
        # the six line passes, two pairs of them concurrent: the step's span on the library stream
        group = ["cfd_step"]
        kname = "cfd_vti_step" if mdl == _lib.VTI else "cfd_step"
    elif mdl == _lib.ISO:
        group = ["iso_step"]
        kname = "iso_step"
    else:
        group = [k for k in ("fd_grad", "fd_div_step") if k in kern]
        kname = "fd_step"
    step_ms = sum(kern[k]["mean_ms"] for k in group if k in kern)
    # PML state bytes per step: 2nd-order form: q1,q2 read+write per face cell (16 B);
    # 1st-order form: 3 derivative components x 2 rounds x (read+write) = 48 B per face cell
    face_cells = 2 * nb * (nx * ny + ny * nz + nx * nz) if nb > 1 else 0
    pml_b = (16 if (eng == _lib.ENGINE_FD and mdl == _lib.ISO) else 48) * face_cells
    alg_bytes = VOL_BYTES[w["model"]] * ncell + pml_b
    achieved = alg_bytes / (step_ms * 1e-3) / 1e9 if step_ms > 0 else 0.0
    traffic = parse_traffic(workload)
    launches = 6 if eng == _lib.ENGINE_CFD and ny > 1 else 3 if eng == _lib.ENGINE_CFD else len(group)
    roofline = {"kernel": kname, "launches_per_step": launches, "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes_per_step": alg_bytes, "step_kernel_ms": round(step_ms, 4)}
    shots_per_hour = 3600.0 / (nt_shot * el / K)
    p.close()
    cpu_line = cpu_baseline(vel, eps, dl, w, eng, mdl, dt, cpu_budget, nt_shot) if cpu else None
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": workload, "grid": [nx, ny, nz], "spacing_m": w["d"], "engine": w["engine"],
                   "order": w["order"], "model": w["model"], "pml": nb, "dt_s": round(dt, 7), "nt_shot": nt_shot,
                   "receivers": int(rcv.shape[0]), "snap_every": w["snap_every"], "snap_format": "fp32",
                   "step": "1 forward + 1 backward RTM time step (injection, snapshot, imaging on cadence)"},
        "shots_per_hour": round(shots_per_hour, 3),
        "roofline": roofline,
        "cpu_baseline": cpu_line,
        "kernels": kern,
    }
    return out


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--workload", default="vti-cfd10-400",
                    choices=sorted(WORKLOADS) + ["tti-fd16-600-shot", "adjacent-400"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-overlap", action="store_true",
                    help="compact engine: run its line passes one after another (for per-kernel counter passes)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--batched-shots", type=int, default=8, help="shots of the batched-vs-plain leg (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=2500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--legs", default="iso-fd8-400,tti-fd16-600",
                    help="further workloads timed after the headline (comma list; empty: none)")
    ap.add_argument("--leg-steps", type=int, default=10)
    ap.add_argument("--convae-steps", type=int, default=4000,
                    help="autoencoder training steps (0: skip that leg; the full schedule is 20000)")
    ap.add_argument("--convae-models", type=int, default=8, help="seeded 256^3 models the autoencoder trains on")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: the hot path runs on a single MI355X "
              f"(requested --gpus {args.gpus}, WORLD_SIZE {world}); sharding shots over several GPUs is a "
              f"design, not something this benchmark runs.", file=sys.stderr)
        sys.exit(2)

    from konglong_amd import _lib

    _lib.lib()
    if _lib.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "tti-fd16-600-shot":
        print(json.dumps(tti_full_shot()))
        return
    if args.workload == "adjacent-400":
        print(json.dumps({"workload": "adjacent-400", **adjacent_physics()}))
        return
    out = measure(args.workload, args.steps, args.warmup, cpu=not args.no_cpu_baseline and
                  int(os.environ.get("RANK", "0")) == 0, cpu_budget=args.cpu_budget, overlap=not args.no_overlap)
    _progress(f"{args.workload}: {out['value']} Gcell-updates/s, frac {out['roofline']['frac']}")
    # the other two plain-FP32 stencil workloads, timed the same way on the same GPU (fewer steps)
    legs = [x for x in args.legs.split(",") if x and x != args.workload]
    if legs:
        out["legs"] = {}
        for name in legs:
            try:
                leg = measure(name, args.leg_steps, 2, cpu=False)
                _progress(f"{name}: {leg['value']} Gcell-updates/s, frac {leg['roofline']['frac']}")
                out["legs"][name] = {k: leg[k] for k in ("value", "unit", "ms_per_step", "shots_per_hour", "roofline",
                                                          "config", "kernels")}
            except Exception as e:  # the headline line is still printed
                out["legs"][name] = {"error": f"{type(e).__name__}: {e}"}
    if args.convae_steps > 0:
        try:
            out["learned_convae"] = convae_vs_plain(args.convae_models, args.convae_steps)
        except Exception as e:  # the headline line is still printed
            out["learned_convae"] = {"error": f"{type(e).__name__}: {e}"}
    if args.batched_shots > 0:
        try:
            out["batched"] = batched_vs_plain(args.batched_shots, args.batched_steps, args.batch)
        except Exception as e:  # the headline line is still printed
            out["batched"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
