"""Receiver-side pre-processing on the GPU: PhaseCorrection's backward time-dispersion correction and
record resampling (bgprtm/PhaseCorrection.cpp, as ExtractReceiverData::resampleRecData uses it)."""
import ctypes as C

import numpy as np

from ._lib import lib, check, DeviceArray


class PhaseCorrection:
    """PhaseCorrection(nt, dt, Dt, t0, fmax, BWD, w_pow, phase_deg, do_tdsp, ntOut, dtOut, t0Out) with the FFT
    length ntfft (default: klrtm_fft_size of rint(1.5 nt), the reference's tpad_ratio 0.5)."""

    def __init__(self, nt, dt, Dt, t0, fmax, ntfft=None, w_pow=0.0, phase_deg=0.0, ntOut=-1, dtOut=-1.0, t0Out=0.0):
        if ntfft is None:
            ntfft = int(lib().klrtm_fft_size(int(round(nt * 1.5))))
        self.h = C.c_void_p()
        check(lib().klrtm_phasecorr_create(C.byref(self.h), nt, dt, Dt, t0, fmax, ntfft, w_pow, phase_deg, 1, ntOut,
                                           dtOut, t0Out))
        n_out, nw1 = C.c_int(), C.c_int()
        check(lib().klrtm_phasecorr_info(self.h, C.byref(n_out), C.byref(nw1)))
        self.nt, self.nt_out, self.nw1, self.ntfft = nt, n_out.value, nw1.value, ntfft

    def backward(self, traces, out=None, scaler=1.0):
        """applyBackward on device traces [ntr][nt] (DeviceArray) into out [ntr][nt_out]; numpy in -> numpy out."""
        host = isinstance(traces, np.ndarray)
        d_in = DeviceArray.from_host(np.ascontiguousarray(traces, np.float32)) if host else traces
        ntr = d_in.shape[0] if len(d_in.shape) > 1 else 1
        d_out = out if out is not None else DeviceArray((ntr, self.nt_out))
        check(lib().klrtm_phasecorr_backward(self.h, d_in.ptr, ntr, d_out.ptr, scaler))
        return d_out.download() if host else d_out

    def close(self):
        if self.h:
            lib().klrtm_phasecorr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def resample_record(record, dt, dt_prop, t0, max_freq, dt_out, nt_out, keep_t0=False, dipole=None):
    """ExtractReceiverData::resampleRecData (bgprtm/ExtractReceiverData.cpp:75-110): the raw record
    [nr][nt] sampled at dt from t0 (<= 0) through PhaseCorrection(nt, dt, dt_prop, -t0, max_freq, BWD, w_pow,
    phase_deg, 1, nt_out, dt_out, keep_t0 ? -t0 : 0), the time-dispersion correction of the propagation step
    dt_prop and the resampling to dt_out.  dipole = (dz, vsurf) for a dipole (true-amplitude) receiver
    spread: w^-1 and a 90 degree phase, scaled by 2 vsurf / dz.  Returns [nr][nt_out] (numpy)."""
    record = np.ascontiguousarray(record, np.float32)
    nr, nt = record.shape
    w_pow, phase_deg, dt_spread = 0.0, 0.0, 1.0
    if dipole is not None:
        dz, vsurf = dipole
        w_pow, phase_deg, dt_spread = -1.0, 90.0, float(np.float32(0.5) * np.float32(dz) / np.float32(vsurf))
    pc = PhaseCorrection(nt, dt, dt_prop, -t0, max_freq, w_pow=w_pow, phase_deg=phase_deg, ntOut=nt_out, dtOut=dt_out,
                         t0Out=-t0 if keep_t0 else 0.0)
    out = pc.backward(record, scaler=float(np.float32(1.0) / np.float32(dt_spread)))
    pc.close()
    return out


def write_record(path, data, dt, t0=0.0, xyz=None):
    """ExtractReceiverData::saveRecFile's content in this build's record file (klrtm_record_write)."""
    d = np.ascontiguousarray(data, np.float32)
    c = None if xyz is None else np.ascontiguousarray(xyz, np.float32).reshape(d.shape[0], 3)
    check(lib().klrtm_record_write(str(path).encode(), d.shape[0], d.shape[1], dt, t0,
                                   None if c is None else c.ctypes.data, d.ctypes.data))


def read_record(path):
    """(data [nr][nt], dt, t0, xyz [nr][3]) from a record file."""
    nr, nt, dt, t0 = C.c_int(), C.c_int(), C.c_float(), C.c_float()
    check(lib().klrtm_record_read_header(str(path).encode(), C.byref(nr), C.byref(nt), C.byref(dt), C.byref(t0)))
    xyz = np.zeros((nr.value, 3), np.float32)
    data = np.zeros((nr.value, nt.value), np.float32)
    check(lib().klrtm_record_read(str(path).encode(), xyz.ctypes.data, data.ctypes.data))
    return data, dt.value, t0.value, xyz
