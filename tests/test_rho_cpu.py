"""Oracle pinning of density (KernelCPU::divideRho and the vel*rho accumulates, bgprtm/KernelCPU.cpp:
93-134, 340-349; bgprtm/FdEngine_fd.cpp:409-464; the force_tti switch, bgprtm/ModelLoader.cpp:63):

* rho = 1 leaves VTI/TTI bit for bit unchanged (x/1 and x*1 are exact), and turns ISO into the
  first-derivative form, which equals VTI with eps = del = 0 bit for bit (that scalar operator
  multiplies by exactly 1);
* a constant density c cancels: rho * div(grad p / rho) is the operator without density, to rounding
  (a dropped division or multiplication leaves a factor c);
* the first-derivative form differs from the Laplacian form (rho really switches ISO over).
"""
import numpy as np
import pytest

from oracle import Engine
from tests.helpers import rel_l2, smooth_field

ISO, VTI, TTI = 1, 2, 3
SHAPE = (14, 15, 16)


def _engine(model, rho=None, order=8, engine=7):
    ny, nx, nz = SHAPE
    e = Engine(nx, ny, nz, 10.0, 10.0, 10.0, engine, order)
    vel = smooth_field(SHAPE, 0.05, 0.15, 71)
    eps = smooth_field(SHAPE, 0.0, 0.2, 72) if model != ISO else None
    dl = smooth_field(SHAPE, 0.0, 0.1, 73) if model != ISO else None
    e.set_model(vel, eps, dl)
    e.set_rho(rho)
    e.set_boundary([4] * 6, 1e-3, 3000.0)
    return e, vel


def _apply(e, model, seed=5):
    rng = np.random.default_rng(seed)
    p0 = rng.standard_normal(SHAPE).astype(np.float32)
    pb = rng.standard_normal(SHAPE).astype(np.float32)
    pr = np.zeros(SHAPE, np.float32)
    before = pb.copy()
    e.clean_pml()
    e.kernel(model, p0, pb, pr)
    return pb, pb.astype(np.float64) - before  # pb - vel*Op(p0), and the increment -vel*Op(p0)


@pytest.mark.parametrize("model,engine", [(VTI, 7), (VTI, 4), (ISO, 4)], ids=["FD-VTI", "CFD-VTI", "CFD-ISO"])
def test_unit_density_is_exact(model, engine):
    a = _apply(_engine(model, engine=engine)[0], model)[0]
    b = _apply(_engine(model, np.ones(SHAPE, np.float32), engine=engine)[0], model)[0]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_iso_with_density_takes_the_first_derivative_form():
    iso_rho, d_rho = _apply(_engine(ISO, np.ones(SHAPE, np.float32))[0], ISO)
    e, vel = _engine(VTI)
    z = np.zeros(SHAPE, np.float32)
    e.set_model(vel, z, z)
    vti0 = _apply(e, VTI)[0]
    assert np.array_equal(iso_rho.view(np.uint32), vti0.view(np.uint32))
    d_lap = _apply(_engine(ISO)[0], ISO)[1]
    assert rel_l2(d_rho, d_lap) > 1e-3


@pytest.mark.parametrize("model", [ISO, VTI])
def test_constant_density_cancels(model):
    ref = _apply(_engine(model, np.ones(SHAPE, np.float32))[0], model)[1]
    got = _apply(_engine(model, np.full(SHAPE, 2.3, np.float32))[0], model)[1]
    assert np.abs(ref).max() > 0 and rel_l2(got, ref) < 1e-5


@pytest.mark.parametrize("model,engine", [(ISO, 7), (VTI, 7), (ISO, 4), (VTI, 4)],
                         ids=["FD-ISO", "FD-VTI", "CFD-ISO", "CFD-VTI"])
def test_power_of_two_density_cancels_exactly(model, engine):
    """rho = 2 scales the gradient by 1/2 and each accumulate by 2 exactly, and the linear solves and
    the scalar operator (a function of ratios) commute with that scaling in floating point: the result
    equals rho = 1 bit for bit.  (The compact scheme amplifies the rounding of other constants near
    Nyquist, so those are only checked on the FD engine above.)"""
    a = _apply(_engine(model, np.ones(SHAPE, np.float32), engine=engine)[0], model)[0]
    b = _apply(_engine(model, np.full(SHAPE, 2.0, np.float32), engine=engine)[0], model)[0]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_variable_density_changes_the_operator():
    rho = smooth_field(SHAPE, 1.0, 2.5, 74)
    a = _apply(_engine(ISO, np.ones(SHAPE, np.float32))[0], ISO)[1]
    b = _apply(_engine(ISO, rho)[0], ISO)[1]
    assert rel_l2(a, b) > 1e-2
