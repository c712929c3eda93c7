"""Oracle pinning of the Q (viscoacoustic) branches of Propagator::Vel2 and Propagator::update2nd
(bgprtm/Propagator.cpp:725-741, 1789-1816) by properties the scheme fixes:

* without attenuation (1/Q = 0) both reduce to the plain Vel2 and leapfrog, bit for bit;
* a spatially constant field is a fixed point once the memory variables hold their steady state
  c/(1 - e) times the field: the (1 + e) and -2 weights cancel the cqsum term exactly, so a dropped
  term, a wrong sign or swapped memory-variable sets break it;
* each memory-variable set advances every other step (the sets swap with the wavefields), so after
  step n the set written then holds the decimated geometric sum c * sum_k e^k w1[n - 2k], which is
  summed here directly in FP64;
* Vel2's correction follows its closed form in FP64.
"""
import numpy as np
import pytest

import oracle


def _coeffs(order, dtype):
    cq0 = 0.04
    cq = [0.3, 0.2, 0.1, 0.05, 0.02][:order]
    wq = [0.9, 0.7, 0.5, 0.3, 0.1][:order]
    return np.array([cq0] + cq + wq, dtype)


@pytest.mark.parametrize("order", [1, 3, 5])
def test_update2nd_q_without_attenuation_is_the_leapfrog(order):
    rng = np.random.default_rng(order)
    shape = (3, 5, 7)
    w0 = rng.standard_normal(shape).astype(np.float32)
    w1 = rng.standard_normal(shape).astype(np.float32)
    d0 = rng.standard_normal((order,) + shape).astype(np.float32)
    d1 = rng.standard_normal((order,) + shape).astype(np.float32)
    want = w0.copy()
    oracle.lib().orc_update2nd_f(want.ctypes.data, w1.ctypes.data, want.size)
    c = _coeffs(order, np.float32)
    d0_before = d0.copy()
    oracle.update2nd_q(w0, w1, np.zeros(shape, np.float32), d0, d1, c)
    assert np.array_equal(w0, want)
    # the memory variables still advance: d0 <- c*w1 + e*d0 (exact: two roundings in FP64 agree)
    for l in range(order):
        ref = (np.float64(c[1 + l]) * w1 + np.float64(c[1 + order + l]) * d0_before[l])
        assert np.allclose(d0[l], ref, rtol=2e-7, atol=1e-7)


# the leapfrog's double root at 1 lets rounding drift grow like steps^2 * eps
@pytest.mark.parametrize("dtype,steps,tol", [(np.float64, 200, 1e-10), (np.float32, 50, 5e-4)])
def test_constant_field_is_a_fixed_point(dtype, steps, tol):
    order = 3
    c = _coeffs(order, dtype)
    cq, wq = c[1:1 + order], c[1 + order:]
    shape = (2, 3, 4)
    C0 = dtype(1.7)
    invq = np.full(shape, 0.05, dtype)
    w0 = np.full(shape, C0, dtype)
    w1 = w0.copy()
    sets = [np.stack([np.full(shape, cq[l] * C0 / (1 - wq[l]), dtype) for l in range(order)]) for _ in range(2)]
    for _ in range(steps):
        oracle.update2nd_q(w0, w1, invq, sets[0], sets[1], c)
        w0, w1 = w1, w0
        sets.reverse()
    assert np.abs(w1 / C0 - 1).max() < tol
    assert np.abs(w0 / C0 - 1).max() < tol


def test_constant_field_drifts_when_a_term_is_wrong():
    """The fixed point is sensitive: with the -2*d1 weight halved (a plausible slip, planted here in
    a numpy restatement, not in the oracle) the field leaves C within a few steps."""
    order = 3
    c = _coeffs(order, np.float64)
    cq, wq = c[1:1 + order], c[1 + order:]
    C0, q = 1.7, 0.05
    w0 = w1 = C0
    A = [cq[l] * C0 / (1 - wq[l]) for l in range(order)]
    B = list(A)
    for _ in range(20):
        u = w1 * (2 + q * cq.sum()) - w0
        for l in range(order):
            u += q * ((1 + wq[l]) * A[l] - B[l] * 1.0)
            A[l] = cq[l] * w1 + wq[l] * A[l]
        w0, w1 = w1, u
        A, B = B, A
    assert abs(w1 / C0 - 1) > 1e-3


def test_memory_variables_are_decimated_geometric_sums():
    order = 2
    c = _coeffs(order, np.float64)
    shape = (1, 2, 3)
    rng = np.random.default_rng(7)
    nt = 41
    # a spatially constant w1 sequence: drive the update with w1[n] and discard w0
    seq = rng.standard_normal(nt)
    sets = [np.zeros((order,) + shape), np.zeros((order,) + shape)]
    invq = np.full(shape, 0.03)
    for n in range(nt):
        w1 = np.full(shape, seq[n])
        w0 = np.zeros(shape)
        oracle.update2nd_q(w0, w1, invq, sets[0], sets[1], c)
        written = sets[0]
        for l in range(order):
            cl, el = c[1 + l], c[1 + order + l]
            want = cl * sum(el ** k * seq[n - 2 * k] for k in range(n // 2 + 1))
            assert np.allclose(written[l], want, rtol=1e-13, atol=1e-15)
        sets.reverse()


def test_vel2_q_closed_form_and_reduction():
    rng = np.random.default_rng(3)
    shape = (3, 4, 5)
    v = rng.uniform(1500, 4500, shape).astype(np.float32)
    dt = np.float32(7e-4)
    # 1/Q = 0: the plain Vel2 (v*v)*dt^2 bit for bit, 1/Q stays 0
    vel, iq = v.copy(), np.zeros(shape, np.float32)
    oracle.vel2_q(vel, iq, dt, _coeffs(2, np.float32))
    assert np.array_equal(vel, (v * v) * (dt * dt)) and not iq.any()
    # attenuating: closed form in FP64
    q0 = rng.uniform(0.0, 0.1, shape)
    vel, iq = v.copy(), q0.astype(np.float32)
    cq0 = 0.04
    oracle.vel2_q(vel, iq, dt, _coeffs(2, np.float32))
    corr = 1 / (1 - cq0 * q0)
    want = v.astype(np.float64) ** 2 * float(dt) ** 2 * (1 - q0 ** 2 / 4 + q0 ** 4 / 48) * corr
    assert np.allclose(vel, want, rtol=3e-6)
    assert np.allclose(iq, q0 * corr, rtol=3e-6)
