"""GPU parity of injection/extraction, snapshot codecs, imaging condition and
the whole-shot driver (Propagator::mod / Propagator::rtm) against the oracle."""
import numpy as np
import pytest

import oracle
from oracle import pipeline
from konglong_amd._lib import DeviceArray, ENGINE_FD, ENGINE_CFD, ISO, VTI, SNAP_FP32, SNAP_C16, SNAP_FP16, SNAP_BF16
from konglong_amd.receiver import Receiver
from konglong_amd.codec import WaveFieldCompress, blockq_encode, blockq_decode, blockq_nscales, nshort_volume
from konglong_amd import imaging
from konglong_amd.propagator import Propagator, shot_config
from tests.helpers import rel_l2, smooth_field

pytestmark = pytest.mark.gpu

SHAPE = (30, 33, 36)  # ny, nx, nz


def _rcv_xyz():
    pts = [[x, y, 12.0] for y in range(0, 30, 3) for x in range(0, 33, 3)]
    pts += [[16.3, 14.7, 20.2], [0.0, 0.0, 12.0], [32.0, 29.0, 12.0], [5.5, 25.25, 12.0]]
    return np.array(pts, np.float32)


def test_receiver_tables_match_oracle(klrtm):
    ny, nx, nz = SHAPE
    xyz = _rcv_xyz()
    for band in ((0.3, 0.3), (0.4, 0.48)):
        r = Receiver(nx, ny, nz, xyz, 5, band[0], band[1], 10, 1e-3)
        t = r.tables()
        o = pipeline.spread_tables(nx, ny, nz, xyz, 5, band[0], band[1])
        for key in ("lx", "ly", "lz", "offx", "offy", "offz"):
            assert np.array_equal(t[key], o[key]), key
        for key in ("kx", "ky", "kz"):
            assert np.array_equal(t[key].view(np.uint32), o[key].view(np.uint32)), key


def test_inject_and_extract_match_oracle(klrtm):
    ny, nx, nz = SHAPE
    xyz = _rcv_xyz()
    nr, nt = xyz.shape[0], 6
    print(f"[inject/extract] nr={nr} grid={SHAPE}", flush=True)
    rng = np.random.default_rng(2)
    traces = rng.standard_normal((nr, nt)).astype(np.float32)
    vel = smooth_field(SHAPE, 0.5, 1.5, 3)
    dt_dxyz = np.float32(1.0 / 1000.0)
    r = Receiver(nx, ny, nz, xyz, 5, 0.3, 0.3, nt, dt_dxyz)
    r.set_traces(traces)
    w0 = np.zeros(SHAPE, np.float32)   # injected increments compared directly
    dw, dv = DeviceArray.from_host(w0), DeviceArray.from_host(vel)
    r.apply(dw, 3, dv, 1.0)
    r.apply(dw, -1, dv, 1.0)   # outside [0, nt): no-op
    r.apply(dw, nt, dv, 1.0)
    got = dw.download()
    t = pipeline.spread_tables(nx, ny, nz, xyz, 5, 0.3, 0.3)
    want = w0.copy()
    pipeline.inject(want, vel, SHAPE, t, traces[:, 3] * dt_dxyz)
    assert np.abs(want).max() > 0
    assert rel_l2(got, want) < 1e-6
    w = rng.standard_normal(SHAPE).astype(np.float32)
    r.extract(DeviceArray.from_host(w), 2)
    rec = r.record()
    assert not rec[:, [0, 1, 3, 4, 5]].any()
    assert rel_l2(rec[:, 2], pipeline.extract(w, SHAPE, t)) < 1e-6
    # one trace: the same products and sums as the reference, bit for bit
    one = Receiver(nx, ny, nz, xyz[-4:-3], 5, 0.3, 0.3, nt, dt_dxyz)
    one.set_traces(traces[-4:-3])
    d1 = DeviceArray((ny, nx, nz))
    one.apply(d1, 3, dv, 1.0)
    w1 = np.zeros(SHAPE, np.float32)
    pipeline.inject(w1, vel, SHAPE, pipeline.spread_tables(nx, ny, nz, xyz[-4:-3], 5, 0.3, 0.3), traces[-4:-3, 3] * dt_dxyz)
    assert np.array_equal(d1.download().view(np.uint32), w1.view(np.uint32))


def test_compress16_bit_exact(klrtm):
    rng = np.random.default_rng(4)
    for shape in ((45, 44, 30), (7, 50, 23), (64, 64, 64)):
        ny, nx, nz = shape
        v = (rng.standard_normal(shape) * np.exp(rng.uniform(-20, 5, shape))).astype(np.float32)
        v[1] = 0.0
        v[2] *= 1e-27          # below the eps threshold -> zero slice
        dv = DeviceArray.from_host(v)
        dc = DeviceArray((nshort_volume(nz, nx, ny),), np.uint16)
        wfc = WaveFieldCompress()
        wfc.compress(dv, dc, nx, ny, nz)
        code = dc.download()
        want = oracle.compress16(v)
        assert np.array_equal(code, want)
        do = DeviceArray(shape)
        wfc.uncompress(dc, do, nx, ny, nz)
        got = do.download()
        ref = oracle.uncompress16(want, ny, nx, nz)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_decode16_every_code(klrtm):
    # one slice holding all 65536 codes with scale 1: decode must equal the table
    nx, nz = 256, 256
    code = np.zeros(nx * nz + 2, np.uint16)
    code[:2] = np.array([1.0], np.float32).view(np.uint16)
    code[2:] = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    dc = DeviceArray.from_host(code)
    do = DeviceArray((1, nx, nz))
    WaveFieldCompress().uncompress(dc, do, nx, 1, nz)
    got = do.download().ravel()
    t = oracle.codec16_table()
    # nx > 40 and nz > 20 apply the edge taper; compare on untapered samples
    ix, iz = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    keep = ((ix >= 20) & (ix < nx - 20) & (iz >= 10) & (iz < nz - 10)).ravel()
    a, b = got[keep], t[keep]
    assert np.array_equal(np.isnan(a), np.isnan(b))
    assert np.array_equal(a[~np.isnan(a)].view(np.uint32), b[~np.isnan(b)].view(np.uint32))


@pytest.mark.parametrize("fmt,bound", [(SNAP_FP16, 2.0 ** -11), (SNAP_BF16, 2.0 ** -8)])
def test_blockq_roundtrip(klrtm, fmt, bound):
    rng = np.random.default_rng(6)
    shape = (19, 21, 70)
    ny, nx, nz = shape
    v = (rng.standard_normal(shape) * np.exp(rng.uniform(-6, 3, shape))).astype(np.float32)
    v[:8, :8, :8] = 0.0
    # a block holding only subnormal-scale residue next to exact zeros: stored as zero, never NaN
    v[8:16, :8, :8] = 0.0
    v[9, 2, 3] = np.float32(3e-39)
    v[10, 4, 5] = np.float32(-1e-36)
    dv = DeviceArray.from_host(v)
    dq = DeviceArray(shape, np.uint16)
    ns = blockq_nscales(nx, ny, nz)
    ds = DeviceArray((ns,))
    blockq_encode(dv, dq, ds, nx, ny, nz, fmt)
    do = DeviceArray(shape)
    blockq_decode(dq, ds, do, nx, ny, nz, fmt)
    out = do.download()
    sc = ds.download()
    assert np.isfinite(out).all()
    assert not out[:8, :8, :8].any()
    if fmt == SNAP_FP16:  # scale 1e-36 * 2^-15 is subnormal: the block is stored as zero
        assert not out[8:16, :8, :8].any()
    else:                 # BF16 keeps scale 1e-36
        assert abs(out[10, 4, 5] - v[10, 4, 5]) <= bound * abs(v[10, 4, 5])
    # per-block max |x| is what the scale records (FP16 keeps a 2^-15 factor); below FLT_MIN it is 0
    bmax = np.zeros(((ny + 7) // 8, (nx + 7) // 8, (nz + 7) // 8), np.float32)
    for by in range(bmax.shape[0]):
        for bx in range(bmax.shape[1]):
            for bz in range(bmax.shape[2]):
                bmax[by, bx, bz] = np.abs(v[by*8:by*8+8, bx*8:bx*8+8, bz*8:bz*8+8]).max()
    fac = 2.0 ** -15 if fmt == SNAP_FP16 else 1.0
    want = (bmax * np.float32(fac)).astype(np.float32)
    want[want < np.float32(1.17549435e-38)] = 0.0
    assert np.array_equal(sc.reshape(bmax.shape), want)
    err = np.abs(out - v)
    rel = err[np.abs(v) > 1e-3 * np.abs(v).max()] / np.abs(v[np.abs(v) > 1e-3 * np.abs(v).max()])
    assert rel.max() <= bound * 1.01
    assert rel_l2(out, v) < bound


def test_ic_correlate_matches_numpy(klrtm):
    rng = np.random.default_rng(8)
    shape = (9, 10, 17)
    s, r = rng.standard_normal(shape).astype(np.float32), rng.standard_normal(shape).astype(np.float32)
    ds, dr = DeviceArray.from_host(s), DeviceArray.from_host(r)
    img, ill = DeviceArray(shape), DeviceArray(shape)
    img64, ill64 = DeviceArray(shape, np.float64), DeviceArray(shape, np.float64)
    imaging.correlate(ds, dr, 10, 9, 17, 0.5, image=img, illum=ill, image64=img64, illum64=ill64)
    tz = np.linspace(0, 1, 6).astype(np.float32)
    dtz = DeviceArray.from_host(tz)
    imaging.correlate(ds, dr, 10, 9, 17, -1.0, taper_z=dtz, ntaper_z=6, image=img, illum=ill, image64=img64,
                      illum64=ill64)
    S1 = s * np.float32(0.5)
    S2 = s.copy()
    S2[:, :, :6] *= tz
    want = (S1 * r) + (S2 * r)
    assert np.array_equal(img.download(), (np.zeros(shape, np.float32) + S1 * r) + S2 * r)
    assert np.allclose(img64.download(), S1.astype(np.float64) * r + S2.astype(np.float64) * r, rtol=1e-12)
    assert np.allclose(ill.download(), S1 * S1 + S2 * S2, rtol=1e-6)
    assert rel_l2(img.download(), want) < 1e-7


# ---------------------------------------------------------------------------
# whole shot
def _shot(engine, order, model, snap_format=SNAP_FP32, nt=160, fuse=1):
    ny, nx, nz = 40, 42, 44
    d, dt = 10.0, 1e-3
    nb = 10
    vel = smooth_field((ny, nx, nz), 1800.0, 2600.0, 21)
    eps = smooth_field((ny, nx, nz), 0.0, 0.2, 22) if model == VTI else None
    dl = smooth_field((ny, nx, nz), 0.0, 0.1, 23) if model == VTI else None
    rcv = np.array([[x, y, 12.0] for y in range(2, ny - 2, 3) for x in range(2, nx - 2, 3)], np.float32)
    src = (21.0, 19.0, 14.0)
    k1 = 0.3 if engine == ENGINE_FD else 0.40
    k2 = 0.3 if engine == ENGINE_FD else 0.48
    cfg = shot_config(nx, ny, nz, d, dt, nt, engine, order, model, nbnd=nb, vmax=2600.0, src_xyz=src, f_peak=15.0,
                      spread=6, khpass=k1, khcut=k2, nrcv=rcv.shape[0], snap_every=2, snap_format=snap_format,
                      rec_through_codec=1 if snap_format == SNAP_C16 else 0, tmute_frac=0.1, fuse_step=fuse)
    p = Propagator(cfg)
    p.set_model(vel, eps, dl)
    p.set_receivers(rcv)
    o = pipeline.Shot(nx, ny, nz, d, dt, nt, engine, order, model, [nb] * 6, 2600.0, src, 15.0, 6, k1, k2, rcv,
                      snap_every=2, snap_format=1 if snap_format == SNAP_C16 else 0, tmute_frac=0.1)
    o.set_model(vel, eps, dl)
    p.set_source_wavelet(o.wavelet)
    return p, o


@pytest.mark.parametrize("case", [(ENGINE_FD, 8, ISO, 1), (ENGINE_FD, 8, ISO, 0), (ENGINE_CFD, 10, VTI, 1),
                                  (ENGINE_FD, 16, VTI, 1)],
                         ids=["FD8-ISO-fused", "FD8-ISO-unfused", "CFD10-VTI", "FD16-VTI"])
def test_modelling_record_matches_oracle(klrtm, case):
    engine, order, model, fuse = case
    print(f"[model] {case}", flush=True)
    p, o = _shot(engine, order, model, fuse=fuse)
    rec = p.model()
    want = o.model()
    err = rel_l2(rec, want)
    print(f"  record rel_l2={err:.2e} |rec|max={np.abs(want).max():.3e}", flush=True)
    assert np.abs(want).max() > 0
    assert err < 1e-4
    last = p.wavefield()
    if engine == ENGINE_CFD:
        # engine and single-trace injection are bit-exact: the forward field is identical
        assert np.array_equal(last.view(np.uint32), o.last.view(np.uint32))
    assert rel_l2(last, o.last) < 1e-4


@pytest.mark.parametrize("case", [(ENGINE_FD, 8, ISO, SNAP_FP32), (ENGINE_FD, 8, ISO, SNAP_C16),
                                  (ENGINE_CFD, 10, VTI, SNAP_FP32)],
                         ids=["FD8-ISO-fp32", "FD8-ISO-c16", "CFD10-VTI-fp32"])
def test_rtm_image_matches_oracle(klrtm, case):
    engine, order, model, fmt = case
    print(f"[rtm] {case}", flush=True)
    p, o = _shot(engine, order, model, snap_format=fmt)
    rec = o.model()
    out = p.rtm(rec)
    img, ill = o.rtm(rec)
    e1, e2 = rel_l2(out["image"], img), rel_l2(out["illum"], ill)
    e3 = rel_l2(out["image64"], img)
    print(f"  image rel_l2={e1:.2e} illum rel_l2={e2:.2e} image64-vs-oracle32={e3:.2e}", flush=True)
    assert np.abs(img).max() > 0
    assert e1 < 1e-4 and e2 < 1e-4 and e3 < 1e-4


@pytest.mark.parametrize("mode", [0, 2, 3], ids=["NONE", "16P", "8P"])
@pytest.mark.parametrize("shape", [(23, 31, 37), (44, 42, 45)], ids=["small", "tapered"])
def test_wavefield_codecs_match_oracle(klrtm, mode, shape):
    """COMPRESS_16P / COMPRESS_8P codes and decodes bit for bit; COMPRESS_NONE decode (edge taper)."""
    from konglong_amd.codec import WaveFieldCompress
    ny, nx, nz = shape
    rng = np.random.default_rng(mode + nz)
    v = (rng.standard_normal(shape) * np.float32(2e-3)).astype(np.float32)
    v[3, 4, :] = 0.25   # a constant trace (differences vanish)
    v[5, 6, :] = 0.0    # a zero trace
    wfc = WaveFieldCompress(compression=mode)
    ns = wfc.nshort_volume(nz, nx, ny)
    d_in, d_code, d_out = DeviceArray.from_host(v), DeviceArray((ns,), np.uint16), DeviceArray(shape)
    wfc.compress(d_in, d_code, nx, ny, nz)
    wfc.uncompress(d_code, d_out, nx, ny, nz)
    got = d_out.download()
    print(f"[codec] mode={mode} shape={shape}: max |decode - input| = {np.abs(got - v).max():.3e}", flush=True)
    if mode == 0:
        assert np.array_equal(d_code.download().view(np.float32).reshape(shape), v)
        # decode = input times the edge taper: compare with the oracle's COMPRESS_16 taper on a flat field
        ones = np.ones(shape, np.float32)
        d1 = DeviceArray.from_host(ones)
        wfc.compress(d1, d_code, nx, ny, nz)
        wfc.uncompress(d_code, d_out, nx, ny, nz)
        taper = d_out.download()
        # taper factors multiply one at a time (y, x, z); a unit field gives their product, so allow 1 ulp
        assert np.allclose(got, v * taper, rtol=2.5e-7, atol=0)
        return
    want_code = oracle.compress_pair(v, mode)
    assert np.array_equal(d_code.download().reshape(want_code.shape), want_code)
    want = oracle.uncompress_pair(want_code, ny, nx, nz, mode)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
