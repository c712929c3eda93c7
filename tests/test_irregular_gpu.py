"""IRREGULAR z grids (FdEngine_fd::setJacob, bgprtm/FdEngine_fd.cpp:92-131) on the GPU against the FP32
oracle: FD steps of every model with PML, the compact engine's jacobz scaling, the z first derivative, and a
whole ISO shot.  Tolerance 1e-4 relative L2 after several steps (the FD kernels contract multiply-adds), as the
regular-grid step tests, and bit for bit on the compact engine."""
import numpy as np
import pytest

import oracle
from konglong_amd._lib import DeviceArray, KlrtmError, ENGINE_FD, ENGINE_CFD, ISO, VTI, TTI
from konglong_amd.kernel import KernelCPU
from tests.helpers import rel_l2, smooth_field, gaussian_blob, tilt_axis

pytestmark = pytest.mark.gpu

NY, NX, NZ = 30, 37, 47
D, DT = 10.0, 1.0e-3
NB = (6, 7, 5, 8, 6, 9)


def _zgrid(nz, a=10.0, b=0.12):
    """Depths growing from a to a + 2 b (nz - 1) per cell, as a velocity-driven IRREGULAR grid does."""
    tau = np.arange(nz, dtype=np.float64)
    return (a * tau + b * tau * tau).astype(np.float32)


def _inputs(model, seed=5):
    shape = (NY, NX, NZ)
    v = smooth_field(shape, 1500.0, 2500.0, seed)
    m = (v * v * np.float32(DT * DT)).astype(np.float32)
    eps = smooth_field(shape, 0.0, 0.3, seed + 10) if model != ISO else None
    dl = smooth_field(shape, 0.0, 0.15, seed + 20) if model != ISO else None
    pjx = pjy = None
    if model == TTI:
        pjx, pjy = tilt_axis(shape, seed + 30)
    w1 = gaussian_blob(NY, NX, NZ, NY / 2, NX / 3, NZ / 2, 3.0)
    w0 = (0.98 * w1).astype(np.float32)
    return m, eps, dl, pjx, pjy, w1, w0


CASES = [(8, ISO), (16, ISO), (8, VTI), (16, TTI)]


@pytest.mark.parametrize("case", CASES, ids=[f"FD{o}-{ {1: 'ISO', 2: 'VTI', 3: 'TTI'}[m]}" for o, m in CASES])
def test_irregular_step_parity(klrtm, case):
    order, model = case
    print(f"[irregular-z step] FD{order} model {model} grid {NY}x{NX}x{NZ}", flush=True)
    m, eps, dl, pjx, pjy, w1, w0 = _inputs(model)
    zg = _zgrid(NZ)
    o = oracle.Engine(NX, NY, NZ, D, D, D, ENGINE_FD, order, np.float32)
    o.set_model(m, eps, dl, pjx, pjy)
    o.set_boundary(NB, DT, 2500.0, dim_pml=3)
    o.set_zgrid(zg)
    k = KernelCPU(NX, NY, NZ, D, D, D, ENGINE_FD, order)
    arrs = [DeviceArray.from_host(a) if a is not None else None for a in (m, dl, eps, pjx, pjy)]
    k.setModel(arrs[0], None, arrs[1], arrs[2], arrs[3], arrs[4])
    k.setBoundary(*NB, DT, 2500.0)
    k.set_zgrid(zg)
    ow1, ow0, owr = w1.copy(), w0.copy(), np.zeros_like(w1)
    g1, g0, gr = DeviceArray.from_host(w1), DeviceArray.from_host(w0), DeviceArray((NY, NX, NZ))
    for _ in range(12):
        o.step(model, ow1, ow0, owr)
        ow0, ow1 = ow1, ow0
        k.step(model, g1, g0, gr)
        g0, g1 = g1, g0
    got = g1.download()
    err = rel_l2(got, ow1)
    # the regular-grid step differs: the metric matters at this stretch
    o2 = oracle.Engine(NX, NY, NZ, D, D, D, ENGINE_FD, order, np.float32)
    o2.set_model(m, eps, dl, pjx, pjy)
    o2.set_boundary(NB, DT, 2500.0, dim_pml=3)
    rw1, rw0, rwr = w1.copy(), w0.copy(), np.zeros_like(w1)
    for _ in range(12):
        o2.step(model, rw1, rw0, rwr)
        rw0, rw1 = rw1, rw0
    print(f"  rel_l2 vs oracle {err:.3e}; irregular vs regular {rel_l2(ow1, rw1):.3e}", flush=True)
    assert np.isfinite(got).all() and err <= 1e-4
    assert rel_l2(ow1, rw1) > 1e-2
    k.set_zgrid(None)   # back to the regular weights: the regular step again
    g1, g0 = DeviceArray.from_host(w1), DeviceArray.from_host(w0)
    k.cleanPMLMemory()
    for _ in range(12):
        k.step(model, g1, g0, gr)
        g0, g1 = g1, g0
    assert rel_l2(g1.download(), rw1) <= 1e-4
    k.close()


CFD_CASES = [(ISO, False), (VTI, False), (VTI, True)]


@pytest.mark.parametrize("case", CFD_CASES, ids=["CFD10-ISO", "CFD10-VTI", "CFD10-VTI-rho"])
def test_irregular_compact_step_bit_exact(klrtm, case):
    """The compact engine scales its z derivative by jacobz (FdEngine_cfd::dz1_2D): steps bit-identical to the
    oracle, as on the regular grid; a uniform zgrid leaves the regular step unchanged bit for bit."""
    model, with_rho = case
    print(f"[irregular-z compact] model {model} rho {with_rho}", flush=True)
    m, eps, dl, pjx, pjy, w1, w0 = _inputs(model, seed=8)
    rho = smooth_field((NY, NX, NZ), 1.0, 2.5, 77) if with_rho else None
    zg = _zgrid(NZ)
    outs = []
    for z in (zg, (D * np.arange(NZ)).astype(np.float32), None):
        o = oracle.Engine(NX, NY, NZ, D, D, D, ENGINE_CFD, 10, np.float32)
        o.set_model(m, eps, dl)
        o.set_rho(rho)
        o.set_boundary(NB, DT, 2500.0, dim_pml=3)
        o.set_zgrid(z)
        k = KernelCPU(NX, NY, NZ, D, D, D, ENGINE_CFD, 10)
        arrs = [DeviceArray.from_host(a) if a is not None else None for a in (m, rho, dl, eps)]
        k.setModel(arrs[0], arrs[1], arrs[2], arrs[3])
        k.setBoundary(*NB, DT, 2500.0)
        k.set_zgrid(z)
        ow1, ow0, owr = w1.copy(), w0.copy(), np.zeros_like(w1)
        g1, g0, gr = DeviceArray.from_host(w1), DeviceArray.from_host(w0), DeviceArray((NY, NX, NZ))
        for _ in range(8):
            o.step(model, ow1, ow0, owr)
            ow0, ow1 = ow1, ow0
            k.step(model, g1, g0, gr)
            g0, g1 = g1, g0
        got = g1.download()
        print(f"  zgrid {'stretched' if z is zg else 'uniform' if z is not None else 'none'}: "
              f"rel_l2 vs oracle {rel_l2(got, ow1):.3e}", flush=True)
        assert np.array_equal(got.view(np.uint32), ow1.view(np.uint32))
        outs.append(got)
        k.close()
    assert np.array_equal(outs[1].view(np.uint32), outs[2].view(np.uint32))
    assert rel_l2(outs[0], outs[2]) > 1e-2


def test_irregular_z_derivatives(klrtm):
    print("[irregular-z deriv1]", flush=True)
    zg = _zgrid(NZ)
    rng = np.random.default_rng(9)
    u = smooth_field((NY, NX, NZ), -1.0, 1.0, 4) + np.float32(0.01) * rng.standard_normal((NY, NX, NZ)).astype(np.float32)
    m = np.full((NY, NX, NZ), 0.5, np.float32)
    o = oracle.Engine(NX, NY, NZ, D, D, D, ENGINE_FD, 8, np.float32)
    o.set_model(m)
    o.set_boundary(NB, DT, 2500.0, dim_pml=3)
    o.set_zgrid(zg)
    want = np.zeros_like(u)
    o.deriv1(0, u, want, 0, 1)
    k = KernelCPU(NX, NY, NZ, D, D, D, ENGINE_FD, 8)
    dm = DeviceArray.from_host(m)
    k.setModel(dm)
    k.setBoundary(*NB, DT, 2500.0)
    k.set_zgrid(zg)
    du, dout = DeviceArray.from_host(u), DeviceArray((NY, NX, NZ))
    from konglong_amd._lib import lib, check
    check(lib().klrtm_kernel_deriv1(k.h, 0, du.ptr, dout.ptr, 0, 1))
    first = dout.download()
    assert rel_l2(first, want) <= 1e-5
    with pytest.raises(KlrtmError):   # a rejected grid leaves the engine on the previous one
        k.set_zgrid(zg[::-1].copy())
    k.cleanPMLMemory()   # the PML state of the first call would change the second
    check(lib().klrtm_kernel_deriv1(k.h, 0, du.ptr, dout.ptr, 0, 1))
    assert np.array_equal(dout.download().view(np.uint32), first.view(np.uint32))
    k.close()
    # the compact engine's z derivative, plain and accumulating (with and without density)
    for rho in (None, smooth_field((NY, NX, NZ), 1.0, 2.0, 78)):
        oc = oracle.Engine(NX, NY, NZ, D, D, D, ENGINE_CFD, 10, np.float32)
        oc.set_model(m)
        oc.set_rho(rho)
        oc.set_boundary(NB, DT, 2500.0, dim_pml=3)
        oc.set_zgrid(zg)
        c = KernelCPU(NX, NY, NZ, D, D, D, ENGINE_CFD, 10)
        drho = DeviceArray.from_host(rho) if rho is not None else None
        c.setModel(dm, drho)
        c.setBoundary(*NB, DT, 2500.0)
        c.set_zgrid(zg)
        for iacc in (0, 1):
            want = np.full_like(u, 0.25)
            oc.deriv1(0, u, want, iacc, -1)
            dout.upload(np.full_like(u, 0.25))
            check(lib().klrtm_kernel_deriv1(c.h, 0, du.ptr, dout.ptr, iacc, -1))
            assert np.array_equal(dout.download().view(np.uint32), want.view(np.uint32)), (iacc, rho is not None)
        c.close()
    with pytest.raises(KlrtmError):   # depths must increase
        k2 = KernelCPU(NX, NY, NZ, D, D, D, ENGINE_FD, 8)
        k2.set_zgrid(zg[::-1].copy())


@pytest.mark.parametrize("case", [(ENGINE_FD, 8, ISO), (ENGINE_CFD, 10, VTI)], ids=["FD8-ISO", "CFD10-VTI"])
def test_irregular_shot_matches_oracle(klrtm, case):
    """A whole shot (modelling, then RTM with snapshots and imaging) on an IRREGULAR grid; on the compact
    engine the forward field stays bit-identical to the oracle's, as on the regular grid."""
    from tests.test_pipeline_gpu import _shot
    engine, order, model = case
    print(f"[irregular-z shot] {case}", flush=True)
    p, o = _shot(engine, order, model)
    zg = _zgrid(p.shape[2], 10.0, 0.1)
    p.set_zgrid(zg)
    o.eng.set_zgrid(zg)
    rec = p.model()
    want = o.model()
    err_rec = rel_l2(rec, want)
    if engine == ENGINE_CFD:
        assert np.array_equal(p.wavefield().view(np.uint32), o.last.view(np.uint32))
    out = p.rtm(want, want=("image", "illum"))
    img, ill = o.rtm(want)
    errs = [err_rec, rel_l2(out["image"], img), rel_l2(out["illum"], ill)]
    print(f"  rel_l2 record/image/illum = {errs}", flush=True)
    assert np.abs(img).max() > 0 and max(errs) < 1e-4
    p.close()
