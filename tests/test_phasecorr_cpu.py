"""Pins of the PhaseCorrection restatement (oracle/phasecorr.py, bgprtm/PhaseCorrection.cpp).

* A direct trigonometric sum, computed without FFTs or the phase table, of what the backward correction
  defines: out(it) = (1/N) sum_t x[t] (1 + 2 sum_{w=1}^{nw1-1} cos(w_w (t dt - t'_it) + (it - it0) dsp_w))
  times the filter/phase factors, at the output times t'_it = it dt_out - (t0Out - t0) after the first input
  sample (the input starts at absolute time -t0, the output at -t0Out), it0 = t0Out / dt_out: it fixes the sign of
  the transform, the DC and scaling conventions, the resampling grid and the time shift.
* Physics: a harmonic oscillator integrated by the leapfrog scheme oscillates at the numerical frequency
  w_n with 2 sin(w_n Dt / 2) = W Dt; the dispersion correction (dsp_sin) maps it onto cos(W t).
* Without dispersion and resampling the correction is the identity up to the Nyquist bin.
"""
import numpy as np
import pytest

from oracle.phasecorr import PhaseCorrection, dsp_sin


def _direct(x, pc, dt, Dt, t0, ntfft, w_pow=0.0, phase_deg=0.0):
    """The backward correction as a direct sum over input samples (FP64, no FFT, no table)."""
    nt = x.size
    n_out = pc.nt_dft
    dt_out = float(pc.dt_dft)
    shift = float(pc.t0Out) - float(t0)
    tp = np.arange(n_out) * dt_out - shift            # output times relative to the input's first sample
    it0 = float(pc.t0Out) / dt_out
    dw = 2 * np.pi / (ntfft * dt)
    t = np.arange(nt) * dt
    out = np.zeros(n_out)
    ph = np.deg2rad(phase_deg)
    for w in range(pc.nw1):
        om = w * dw
        dsp = 0.0 if Dt == 0 else dsp_sin(om * Dt) * dt_out / Dt
        amp = 1.0 if w_pow == 0 else (0.0 if w == 0 else om ** w_pow)
        wgt = 1.0 if w == 0 else 2.0
        arg = om * (t[None, :] - tp[:, None]) + (np.arange(n_out)[:, None] - it0) * dsp - ph
        out += wgt * amp * (np.cos(arg) @ x)
    return out / ntfft


@pytest.mark.parametrize("case", [
    dict(Dt=0.0, t0=0.0),
    dict(Dt=0.004, t0=0.0),
    dict(Dt=0.004, t0=-0.05),
    dict(Dt=0.004, t0=-0.05, ntOut=70, dtOut=0.008, t0Out=0.0),
    dict(Dt=0.002, t0=0.0, ntOut=300, dtOut=0.002, t0Out=0.01),
    dict(Dt=0.004, t0=0.0, w_pow=-1.0, phase_deg=90.0),
])
def test_backward_equals_direct_sum(case):
    nt, dt, fmax = 160, 0.004, 60.0
    ntfft = 240
    rng = np.random.default_rng(3)
    x = rng.standard_normal(nt)
    kw = {k: case[k] for k in ("w_pow", "phase_deg", "ntOut", "dtOut", "t0Out") if k in case}
    pc = PhaseCorrection(nt, dt, case["Dt"], case["t0"], fmax, ntfft, **kw)
    got = pc.backward(x)[0]
    want = _direct(x, pc, dt, case["Dt"], case["t0"], ntfft, case.get("w_pow", 0.0), case.get("phase_deg", 0.0))
    # the table's phases are FP32 (angles up to a few hundred radians when resampling): 5e-5 of the output's scale
    assert np.abs(got - want).max() < 5e-5 * np.abs(want).max()


def test_identity_without_dispersion():
    nt, dt, ntfft = 200, 0.002, 300
    x = np.random.default_rng(5).standard_normal(nt)
    pc = PhaseCorrection(nt, dt, 0.0, 0.0, 0.5 / dt, ntfft)
    assert pc.nw1 == ntfft // 2 + 1
    got = pc.backward(x)[0]
    # the Nyquist bin takes weight 2 instead of 1: the remainder is its term (1/N) X_N/2 (-1)^t
    xn = np.sum(np.concatenate([x, np.zeros(ntfft - nt)]) * (-1.0) ** np.arange(ntfft))
    want = x + xn / ntfft * (-1.0) ** np.arange(nt)
    assert np.abs(got - want).max() < 1e-5 * np.abs(x).max()


def test_dispersion_correction_recovers_the_exact_oscillator():
    """Leapfrog u[n+1] = 2u[n] - u[n-1] - (W Dt)^2 u[n] oscillates as cos(w_n t); the correction gives cos(W t)
    (away from the trace ends, whose truncation the frequency-dependent correction smears)."""
    Dt, W, nt = 0.002, 2 * np.pi * 40.0, 1500
    u = np.zeros(nt)
    u[0] = 1.0
    u[1] = 1.0 - 0.5 * (W * Dt) ** 2          # the leapfrog start of a cosine
    for n in range(1, nt - 1):
        u[n + 1] = 2 * u[n] - u[n - 1] - (W * Dt) ** 2 * u[n]
    wn = 2.0 / Dt * np.arcsin(W * Dt / 2)
    t = np.arange(nt) * Dt
    assert np.abs(u - np.cos(wn * t)).max() < 1e-6     # the numerical frequency
    pc = PhaseCorrection(nt, Dt, Dt, 0.0, 120.0, 3000)
    got = pc.backward(u)[0]
    mid = slice(200, 800)
    err = np.abs(got[mid] - np.cos(W * t)[mid]).max()
    assert err < 5e-3
    assert np.abs(got[mid] - np.cos(wn * t)[mid]).max() > 1.0    # the dispersion was there to correct
