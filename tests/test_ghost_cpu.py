"""Oracle pinning of the ghost mirrors of Receiver::apply and Receiver::extract_value
(bgprtm/Receiver.cpp:394-426, 457-507) by properties the mirror fixes:

* injected into zeros with vel = 1, the ghost field is exactly antisymmetric about the mirror plane and
  zero on it (the direct term and its mirror land in the same cell there): an off-by-one in the mirror
  index or a wrong sign breaks this;
* injection and extraction with the same mirror are adjoint (<inject(d), w> = <d, extract(w)>);
* iz_mirror = -1 is the plain injection and extraction, bit for bit.
"""
import numpy as np
import pytest

from oracle import pipeline


def _tables(shape, xyz, spread=5):
    ny, nx, nz = shape
    return pipeline.spread_tables(nx, ny, nz, np.asarray(xyz, np.float32).reshape(-1, 3), spread, 0.3, 0.3)


@pytest.mark.parametrize("zs", [14.0, 15.4, 18.7])
def test_ghost_injection_is_antisymmetric_about_the_mirror(zs):
    shape = (9, 11, 40)
    izm = 12
    t = _tables(shape, [(5.3, 4.6, zs)])
    w = np.zeros(shape, np.float32)
    pipeline.inject(w, np.ones(shape, np.float32), shape, t, np.array([0.7], np.float32), iz_mirror=izm)
    assert np.abs(w).max() > 0
    assert not w[:, :, izm].any()
    for d in range(1, izm + 1):
        assert np.array_equal(w[:, :, izm + d], -w[:, :, izm - d])
    # the direct part sits where the plain injection puts it, below the mirrored band
    plain = np.zeros(shape, np.float32)
    pipeline.inject(plain, np.ones(shape, np.float32), shape, t, np.array([0.7], np.float32))
    assert np.array_equal(w[:, :, 2 * izm + 1:], plain[:, :, 2 * izm + 1:])


def test_ghost_injection_and_extraction_are_adjoint():
    shape = (12, 13, 36)
    izm = 10
    rng = np.random.default_rng(11)
    xyz = [(rng.uniform(3, 9), rng.uniform(3, 8), rng.uniform(11, 16)) for _ in range(7)]
    t = _tables(shape, xyz)
    d = rng.standard_normal(len(xyz)).astype(np.float32)
    w = rng.standard_normal(shape).astype(np.float32)
    inj = np.zeros(shape, np.float32)
    pipeline.inject(inj, np.ones(shape, np.float32), shape, t, d, iz_mirror=izm)
    ext = pipeline.extract(w, shape, t, iz_mirror=izm)
    lhs = float(np.dot(inj.ravel().astype(np.float64), w.ravel()))
    rhs = float(np.dot(d.astype(np.float64), ext))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs))
    # and the mirror matters: without it the extraction differs
    assert np.abs(pipeline.extract(w, shape, t) - ext).max() > 1e-3


def test_no_mirror_is_the_plain_operator():
    shape = (6, 7, 30)
    rng = np.random.default_rng(5)
    t = _tables(shape, [(3.2, 2.9, 12.5), (2.1, 3.3, 14.0)])
    v = rng.uniform(0.5, 1.5, shape).astype(np.float32)
    d = np.array([0.3, -1.1], np.float32)
    a = rng.standard_normal(shape).astype(np.float32)
    b = a.copy()
    pipeline.inject(a, v, shape, t, d)
    pipeline.inject(b, v, shape, t, d, iz_mirror=-1)
    assert np.array_equal(a, b)
    assert np.array_equal(pipeline.extract(a, shape, t), pipeline.extract(a, shape, t, iz_mirror=-1))
