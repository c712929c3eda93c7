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
