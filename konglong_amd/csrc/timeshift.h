// timeshift.h -- FFT time shift of traces (FreqUtil) on the device (internal).
#pragma once

namespace klrtm {
// smallest 2^a 3^b 5^c >= n
int fft_size(int n);
// out[s][0..nt) = in shifted by shift_host[s] samples (circular over ntfft); in and out on the device
int timeshift(const float *in, int nt, int ntfft, int ntr, const float *shift_host, float *out);
}  // namespace klrtm
