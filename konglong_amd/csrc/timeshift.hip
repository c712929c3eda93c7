// timeshift.hip -- FreqUtil::timeshift_in/out (bgprtm/FreqUtil.cpp:24-39) on the device with hipFFT:
// a trace, zero-padded to ntfft, goes to the frequency domain once (R2C); each output trace is that
// spectrum times exp(-i w t/dt) with w = iw * 2 pi / ntfft, back through C2R and scaled by 1/ntfft,
// keeping the first nt samples.  This is the delayed source array of Propagator::setupReceiver
// (bgprtm/Propagator.cpp:783-811): one wavelet, one delay per source.  The shift is circular over
// ntfft samples.  ntfft comes from klrtm_fft_size (the smallest 2^a 3^b 5^c >= n); the reference's
// libCommon::padfft is unvendored.
#include "klrtm_internal.h"
#include "timeshift.h"
#include <hipfft/hipfft.h>
#include <vector>

namespace klrtm {

#define KL_FFT(call)                                                                        \
  do {                                                                                      \
    hipfftResult r_ = (call);                                                               \
    if(r_ != HIPFFT_SUCCESS) {                                                              \
      ::klrtm::set_error("%s failed: hipfft status %d (%s:%d)", #call, (int)r_, __FILE__, __LINE__); \
      rc = KLRTM_EDEVICE;                                                                   \
      goto done;                                                                            \
    }                                                                                       \
  } while(0)

// spectra[s][iw] = W[iw] * exp(-i iw (t_s * 2 pi / ntfft)), the phase formed in FP32 as the reference
__global__ void timeshift_phase_kernel(const float2 *__restrict__ W, float2 *__restrict__ S, const float *__restrict__ tw,
                                       int nw, int ntr) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= (long)nw * ntr) return;
  const int iw = (int)(k % nw), s = (int)(k / nw);
  float sn, cs;
  sincosf(-(float)iw * tw[s], &sn, &cs);
  const float2 w = W[iw];
  S[k] = make_float2(w.x * cs - w.y * sn, w.x * sn + w.y * cs);
}
// out[s][it] = real[s][it] / ntfft for it < nt
__global__ void timeshift_scale_kernel(const float *__restrict__ real, float *__restrict__ out, int nt, int ntfft, int ntr,
                                       float scaler) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= (long)nt * ntr) return;
  const int it = (int)(k % nt), s = (int)(k / nt);
  out[k] = real[(size_t)s * ntfft + it] * scaler;
}

int fft_size(int n) {
  if(n <= 1) return 1;
  for(int m = n;; m++) {
    int r = m;
    for(int f : {2, 3, 5})
      while(r % f == 0) r /= f;
    if(r == 1) return m;
  }
}

int timeshift(const float *in, int nt, int ntfft, int ntr, const float *shift_host, float *out) {
  KL_REQUIRE(in && out && shift_host, KLRTM_EINVAL, "null argument");
  KL_REQUIRE(nt > 0 && ntr > 0 && ntfft >= nt, KLRTM_EINVAL, "need nt > 0, ntr > 0 and ntfft >= nt");
  const int nw = ntfft / 2 + 1;
  int rc = KLRTM_OK;
  float *pad = nullptr, *real = nullptr, *tw = nullptr;
  float2 *W = nullptr, *S = nullptr;
  hipfftHandle fwd = 0, bwd = 0;
  bool have_fwd = false, have_bwd = false;
  std::vector<float> t(ntr);
  // t_over_dt *= dwdt with dwdt = 2 pi / ntfft, in FP32 (bgprtm/FreqUtil.cpp:30-32)
  const float dwdt = 2.0f * (float)M_PI / ntfft;
  for(int s = 0; s < ntr; s++) t[s] = shift_host[s] * dwdt;
  if(hipMalloc(&pad, sizeof(float) * ntfft) != hipSuccess || hipMalloc(&W, sizeof(float2) * nw) != hipSuccess ||
     hipMalloc(&S, sizeof(float2) * (size_t)nw * ntr) != hipSuccess ||
     hipMalloc(&real, sizeof(float) * (size_t)ntfft * ntr) != hipSuccess || hipMalloc(&tw, sizeof(float) * ntr) != hipSuccess) {
    set_error("time-shift work space (%d traces of %d samples) does not fit", ntr, ntfft);
    rc = KLRTM_ENOMEM;
    goto done;
  }
  if(hipMemsetAsync(pad, 0, sizeof(float) * ntfft, stream()) != hipSuccess ||
     hipMemcpyAsync(pad, in, sizeof(float) * nt, hipMemcpyDeviceToDevice, stream()) != hipSuccess ||
     hipMemcpyAsync(tw, t.data(), sizeof(float) * ntr, hipMemcpyHostToDevice, stream()) != hipSuccess) {
    set_error("time-shift staging failed");
    rc = KLRTM_EDEVICE;
    goto done;
  }
  KL_FFT(hipfftPlan1d(&fwd, ntfft, HIPFFT_R2C, 1));
  have_fwd = true;
  KL_FFT(hipfftSetStream(fwd, stream()));
  KL_FFT(hipfftExecR2C(fwd, pad, (hipfftComplex *)W));
  hipLaunchKernelGGL(timeshift_phase_kernel, dim3((unsigned)div_up((long)nw * ntr, 256)), dim3(256), 0, stream(), W, S, tw,
                     nw, ntr);
  if((rc = check_launch("timeshift_phase_kernel"))) goto done;
  {
    int n[1] = {ntfft};
    KL_FFT(hipfftPlanMany(&bwd, 1, n, nullptr, 1, nw, nullptr, 1, ntfft, HIPFFT_C2R, ntr));
  }
  have_bwd = true;
  KL_FFT(hipfftSetStream(bwd, stream()));
  KL_FFT(hipfftExecC2R(bwd, (hipfftComplex *)S, real));
  hipLaunchKernelGGL(timeshift_scale_kernel, dim3((unsigned)div_up((long)nt * ntr, 256)), dim3(256), 0, stream(), real, out,
                     nt, ntfft, ntr, 1.0f / ntfft);
  if((rc = check_launch("timeshift_scale_kernel"))) goto done;
  if(hipStreamSynchronize(stream()) != hipSuccess) {
    set_error("time shift failed on the device");
    rc = KLRTM_EDEVICE;
  }
done:
  if(have_fwd) (void)hipfftDestroy(fwd);
  if(have_bwd) (void)hipfftDestroy(bwd);
  for(void *p : {(void *)pad, (void *)W, (void *)S, (void *)real, (void *)tw})
    if(p) (void)hipFree(p);
  return rc;
}

}  // namespace klrtm

extern "C" {
int klrtm_fft_size(int n) { return klrtm::fft_size(n); }
int klrtm_timeshift(const float *in, int nt, int ntfft, int ntr, const float *shift_over_dt, float *out) {
  return klrtm::timeshift(in, nt, ntfft, ntr, shift_over_dt, out);
}
}
