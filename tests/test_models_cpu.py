"""The CFL time step of ModelPrepare::calcDeltaTime (bgprtm/ModelPrepare.cpp:647-870) on small models:
the TTI branch against the VTI and ISO branches where they must agree, and its dependence on the tilt."""
import numpy as np

from konglong_amd import models
from konglong_amd._lib import ENGINE_FD, ENGINE_CFD, ISO, VTI, TTI


def _fields(shape, seed):
    vel, eps, dl = models.smooth_vti_model(shape, 3.0, seed)
    pjx, pjy = models.tilt_axes(shape, 3.0, seed + 10)
    return vel, eps, dl, pjx, pjy


def test_tti_without_tilt_is_vti():
    vel, eps, dl, _, _ = _fields((6, 7, 8), 5)
    z = np.zeros_like(vel)
    for eng, order in ((ENGINE_FD, 8), (ENGINE_FD, 16), (ENGINE_CFD, 10)):
        a = models.cfl_dt(vel, 12.5, eng, order, TTI, eps, dl, z, z)
        b = models.cfl_dt(vel, 12.5, eng, order, VTI, eps, dl)
        assert abs(a - b) <= 1e-12 * b


def test_tti_isotropic_is_iso_for_any_tilt():
    # with eps = delta = 0 the rotated bound is v*|k| whatever the axis: the ISO bound
    vel, _, _, pjx, pjy = _fields((5, 6, 7), 9)
    z = np.zeros_like(vel)
    for d in (10.0, (10.0, 12.5, 7.5)):
        a = models.cfl_dt(vel, d, ENGINE_FD, 8, TTI, z, z, pjx, pjy)
        b = models.cfl_dt(vel, d, ENGINE_FD, 8, ISO)
        assert abs(a - b) <= 1e-9 * b


def test_tti_bound_follows_the_axis():
    # constant medium, anisotropic: a vertical axis on a grid fine in z gives the VTI step; tilting the
    # axis 90 degrees into x swaps the roles of x and z (the fast direction moves onto the fine axis)
    shape = (3, 3, 3)
    vel = np.full(shape, 3000.0, np.float32)
    eps = np.full(shape, 0.3, np.float32)
    dl = np.full(shape, 0.1, np.float32)
    d = (20.0, 20.0, 5.0)
    z = np.zeros(shape, np.float32)
    vertical = models.cfl_dt(vel, d, ENGINE_FD, 8, TTI, eps, dl, z, z)
    assert abs(vertical - models.cfl_dt(vel, d, ENGINE_FD, 8, VTI, eps, dl)) <= 1e-12 * vertical
    tilted = models.cfl_dt(vel, d, ENGINE_FD, 8, TTI, eps, dl, np.full(shape, -1.0, np.float32), z)
    assert tilted < vertical
    # the tilted case equals a VTI medium on the grid with x and z spacings exchanged
    swapped = models.cfl_dt(vel, (5.0, 20.0, 20.0), ENGINE_FD, 8, VTI, eps, dl)
    assert abs(tilted - swapped) <= 1e-6 * swapped
