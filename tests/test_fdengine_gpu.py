"""GPU parity of the FdEngine and PML class operations (the layer under KernelCPU)
against the oracle: whole-volume Laplacian and single-axis first derivatives with
their PML (FD at 1e-5 relative L2 -- the FD kernels contract multiply-adds; CFD
bit for bit), and the PML class's line rounds bit for bit."""
import numpy as np
import pytest

import oracle
from konglong_amd import _lib
from konglong_amd._lib import DeviceArray, ENGINE_FD, ENGINE_CFD, check, lib
from konglong_amd.kernel import FdEngine_fd, FdEngine_cfd
from tests.helpers import rel_l2, smooth_field

pytestmark = pytest.mark.gpu

NY, NX, NZ = 22, 31, 37
D, DT, VMAX = 10.0, 1.0e-3, 2500.0
NB = (6, 7, 5, 8, 6, 9)


def _pair(engine, order):
    shape = (NY, NX, NZ)
    v = smooth_field(shape, 1500.0, 2500.0, 5)
    m = (v * v * np.float32(DT * DT)).astype(np.float32)
    o = oracle.Engine(NX, NY, NZ, D, D, D, engine, order, np.float32)
    o.set_model(m)
    o.set_boundary(NB, DT, VMAX, dim_pml=3)
    cls = FdEngine_fd if engine == ENGINE_FD else FdEngine_cfd
    g = cls(NX, NY, NZ, D, D, D, order)
    dm = DeviceArray.from_host(m)
    g.setModel(dm)
    g.setBoundary(*NB, DT, VMAX)
    return o, g, dm, m


def test_laplacian(klrtm):
    o, g, _, _ = _pair(ENGINE_FD, 8)
    rng = np.random.default_rng(1)
    for call in range(3):  # the PML state carries over between calls
        u = smooth_field((NY, NX, NZ), -1.0, 1.0, 10 + call) + np.float32(0.01) * rng.standard_normal(
            (NY, NX, NZ)).astype(np.float32)
        want = np.zeros_like(u)
        o.laplacian(u, want)
        du, dout = DeviceArray.from_host(u), DeviceArray((NY, NX, NZ))
        g.laplacian(du, dout)
        err = rel_l2(dout.download(), want)
        print(f"[fdengine] laplacian call {call}: rel_l2 = {err:.3e}", flush=True)
        assert err <= 1e-5


@pytest.mark.parametrize("engine", [ENGINE_FD, ENGINE_CFD], ids=["FD8", "CFD10"])
@pytest.mark.parametrize("axis", [0, 1, 2], ids=["z", "x", "y"])
def test_deriv1(klrtm, engine, axis):
    order = 8 if engine == ENGINE_FD else 10
    o, g, _, m = _pair(engine, order)
    rng = np.random.default_rng(axis)
    for call, (iacc, isign) in enumerate([(0, 1), (1, -1), (0, -1), (1, 1), (0, 1)]):
        u = smooth_field((NY, NX, NZ), -1.0, 1.0, 20 + call) + np.float32(0.01) * rng.standard_normal(
            (NY, NX, NZ)).astype(np.float32)
        base = rng.standard_normal((NY, NX, NZ)).astype(np.float32) if iacc else np.zeros((NY, NX, NZ), np.float32)
        want = base.copy()
        o.deriv1(axis, u, want, iacc, isign)
        du, dout = DeviceArray.from_host(u), DeviceArray.from_host(base)
        g.deriv1(axis, du, dout, iacc, isign)
        got = dout.download()
        err = rel_l2(got, want)
        print(f"[fdengine] {'FD8' if engine == ENGINE_FD else 'CFD10'} axis={axis} iacc={iacc} isign={isign}: "
              f"rel_l2 = {err:.3e}", flush=True)
        if engine == ENGINE_CFD:
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), err
        else:
            assert err <= 1e-5


def test_pml_class_rounds(klrtm):
    """PML::apply2_single over a batch of lines and apply_single_trans, bit for bit."""
    npml, nbatch, n = 12, 50, 40
    coef, feoc = oracle.pml_profile(npml, 12.5, 1e-3, 4000.0)
    gcoef = np.zeros(npml, np.float32)
    gfeoc = np.zeros(npml, np.float32)
    check(lib().klrtm_pml_profile(npml, 12.5, 1e-3, 4000.0, 1e-3, gcoef.ctypes.data_as(_lib._pf),
                                  gfeoc.ctypes.data_as(_lib._pf)))
    assert np.array_equal(gcoef, coef) and np.array_equal(gfeoc, feoc)
    rng = np.random.default_rng(4)
    p = rng.standard_normal((nbatch, n)).astype(np.float32)
    q = rng.standard_normal((nbatch, npml)).astype(np.float32) * np.float32(0.1)
    q2 = rng.standard_normal((nbatch, npml)).astype(np.float32) * np.float32(0.1)
    # bottom layer with symmetry 1 (topbot 2): the layer ends one sample early and p[npml] mirrors
    off = n - npml - 1
    wp, wq, wq2 = p.copy(), q.copy(), q2.copy()
    for i in range(npml):  # column i of every line: the same g, two rounds
        a, wq[:, i] = oracle.pml_round(feoc[i], wp[:, off + i], wq[:, i])
        wp[:, off + i], wq2[:, i] = oracle.pml_round(feoc[i], a, wq2[:, i])
    wp[:, off + npml] = wp[:, off + npml - 1]
    dp, dq, dq2 = DeviceArray.from_host(p), DeviceArray.from_host(q), DeviceArray.from_host(q2)
    dc = DeviceArray.from_host(feoc)
    check(lib().klrtm_pml_apply_lines(dc.ptr, npml, nbatch, n, 1.0, dp.ptr + 4 * off, dq.ptr, dq2.ptr, 1))
    assert np.array_equal(dp.download(), wp)
    assert np.array_equal(dq.download(), wq) and np.array_equal(dq2.download(), wq2)
    # apply_single_trans across a z run
    g = np.float32(coef[3] * np.float32(0.3))
    pz = rng.standard_normal(300).astype(np.float32)
    qz = rng.standard_normal(300).astype(np.float32) * np.float32(0.1)
    wp, wq = oracle.pml_round(g, pz, qz)
    dp, dq = DeviceArray.from_host(pz), DeviceArray.from_host(qz)
    check(lib().klrtm_pml_apply_trans(g, dp.ptr, dq.ptr, None, 300))
    assert np.array_equal(dp.download(), wp) and np.array_equal(dq.download(), wq)
