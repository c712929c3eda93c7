// phasecorr.hip -- PhaseCorrection's backward correction on the device
// (bgprtm/PhaseCorrection.cpp:12-192, the BWD / fft_dft path used to resample and correct the
// time dispersion of recorded traces, ExtractReceiverData::resampleRecData,
// bgprtm/ExtractReceiverData.cpp:75-110).
//
// Per trace: zero-pad to ntfft, R2C (hipFFT, batched), X[0] real, X[iw] *= exp(-i phase)
// (iw dw)^w_pow / ntfft, X[0] halved; then out[it] = sum_iw Re(X[iw] table[iw][it]) with the
// reference's phase table (dispersion phase, resampling grid and time shift).  That sum is a
// real matrix product: out[ntr][nt_dft] = [Re X, -Im X][ntr][2 nw1] x [Re T; Im T][2 nw1][nt_dft],
// run as one rocBLAS SGEMM per chunk of traces.  The table is built on the host in the
// reference's FP32 arithmetic.  The unvendored libfftv forward transform is taken with the
// positive exponent (X = conj of hipFFT's R2C): the convention for which the table reproduces
// the trace when Dt = 0 without resampling.
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include <hipfft/hipfft.h>
#include <rocblas/rocblas.h>
#include <cmath>
#include <complex>
#include <vector>

namespace klrtm {

struct PhaseCorr {
  int nt = 0, ntfft = 0, nw = 0, nw1 = 0, nt_dft = 0, chunk = 0;
  bool resample = false, tdsp = true;
  float *d_tab = nullptr;   // [2 nw1][nt_dft]: Re(table) rows, then Im(table) rows
  float2 *d_fac = nullptr;  // [nw1]: exp(-i phase) (iw dw)^w_pow / ntfft, as the reference forms it
  float *d_pad = nullptr;   // [chunk][ntfft]
  float2 *d_spec = nullptr; // [chunk][nw]
  float *d_a = nullptr;     // [chunk][2 nw1]
  hipfftHandle plan = 0;
  bool have_plan = false;
  rocblas_handle blas = nullptr;
  ~PhaseCorr() {
    if(have_plan) (void)hipfftDestroy(plan);
    if(blas) (void)rocblas_destroy_handle(blas);
    for(void *p : {(void *)d_tab, (void *)d_fac, (void *)d_pad, (void *)d_spec, (void *)d_a})
      if(p) (void)hipFree(p);
  }
};

static inline float round_down_angle(float angle) {  // bgprtm/PhaseCorrection.cpp:47-50
  const float two_pi = 2.0f * (float)M_PI;
  int n = (int)nearbyintf(angle / two_pi);
  return angle - n * two_pi;
}
static inline double dsp_acos(double x) { return acos(1 - 0.5 * x * x) - x; }
static inline double dsp_sin(double x) { return x - 2. * sin(x / 2.0); }

// zero-padded traces for the batched R2C
__global__ void pc_pad_kernel(const float *__restrict__ in, float *__restrict__ pad, int nt, int ntfft, int ntr) {
  const size_t tot = (size_t)ntr * ntfft;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const size_t tr = e / ntfft;
    const int t = (int)(e - tr * ntfft);
    pad[e] = t < nt ? in[tr * nt + t] : 0.f;
  }
}

// A[tr][iw] = Re X, A[tr][nw1 + iw] = -Im X with X = conj(spectrum) * fac (X[0] real before and after,
// then halved), bgprtm/PhaseCorrection.cpp:150-170
__global__ void pc_spectrum_kernel(const float2 *__restrict__ spec, const float2 *__restrict__ fac, float *__restrict__ a,
                                   int nw, int nw1, int ntr) {
  const size_t tot = (size_t)ntr * nw1;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const size_t tr = e / nw1;
    const int iw = (int)(e - tr * nw1);
    const float2 s = spec[tr * nw + iw];
    float xr = s.x, xi = -s.y;
    if(iw == 0) xi = 0.f;
    const float2 f = fac[iw];
    float yr = xr * f.x - xi * f.y, yi = xr * f.y + xi * f.x;
    if(iw == 0) yr = yr * 0.5f, yi = 0.f;
    a[tr * (size_t)(2 * nw1) + iw] = yr;
    a[tr * (size_t)(2 * nw1) + nw1 + iw] = -yi;
  }
}

__global__ void pc_scale_kernel(float *__restrict__ out, size_t n, float s) {
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) out[e] *= s;
}

static unsigned nblk(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 256 * 64)); }

int phasecorr_create(PhaseCorr **out, int nt, float dt, float Dt, float t0, float fmax, int ntfft, float w_pow,
                     float phase_deg, int do_tdsp, int ntOut, float dtOut, float t0Out) {
  KL_REQUIRE(out, KLRTM_EINVAL, "null output");
  KL_REQUIRE(nt > 0 && dt > 0 && fmax > 0 && ntfft >= nt, KLRTM_EINVAL, "need nt > 0, dt > 0, fmax > 0, ntfft >= nt");
  KL_REQUIRE(do_tdsp, KLRTM_EINVAL, "only the tdsp (table) form of the backward correction is built");
  auto *P = new PhaseCorr();
  P->nt = nt, P->ntfft = ntfft, P->nw = ntfft / 2 + 1;
  // constructor (bgprtm/PhaseCorrection.cpp:12-40)
  const float phase = phase_deg * (float)M_PI / 180;
  const std::complex<float> c_phase = std::exp(std::complex<float>(0.0f, -phase));
  P->resample = dtOut > 0 && ntOut > 0;
  float dt_dft, t_shift;
  if(P->resample) P->nt_dft = ntOut, dt_dft = dtOut, t_shift = t0Out - t0;
  else P->nt_dft = nt, dt_dft = dt, t_shift = 0, t0Out = t0 + t_shift;
  const int it_shift = int(t_shift / dt_dft);
  const float t_shift_frac = t_shift - it_shift * dt_dft;
  // init (bgprtm/PhaseCorrection.cpp:60-108), BWD: do_fft_dft, no Jacobian, tdsp_acos_ratio 0
  const float ratio = 0.0f;
  const float it0 = t0Out / dt_dft;
  const double dw = 2.0 * M_PI / (ntfft * dt);
  int nw1 = (int)nearbyintf((float)(2.0f * (float)M_PI * fmax / dw)) + 1;
  nw1 = std::min(P->nw, nw1);
  P->nw1 = nw1;
  const int ndft = P->nt_dft;
  std::vector<float> tab((size_t)2 * nw1 * ndft);
  for(int iw = 0; iw < nw1; iw++) {
    const double x = (double)iw * dw * Dt;
    const float factor = 2.0f;
    const float dsp = Dt == 0 ? 0 : (ratio * dsp_acos(x) + (1 - ratio) * dsp_sin(x)) * dt_dft / Dt;
    for(int it = 0; it < ndft; it++) {
      float phs;
      if(dt == dt_dft) {
        phs = (it - it0) * dsp - ((long(it - it_shift) * iw) % ntfft) * dw * dt_dft + t_shift_frac * iw * dw;
      } else {
        phs = (it - it0) * dsp - (float)dw * dt_dft * it * iw + (float)dw * t_shift * iw;
        phs = round_down_angle(phs);
      }
      const std::complex<float> c = std::exp(std::complex<float>(0.0f, phs)) * factor;
      tab[(size_t)iw * ndft + it] = c.real();
      tab[(size_t)(nw1 + iw) * ndft + it] = c.imag();
    }
  }
  // fft_dft's per-frequency factor (bgprtm/PhaseCorrection.cpp:150-158)
  std::vector<float2> fac(nw1);
  const float dwf = 2.0f * M_PI / (ntfft * dt);
  const float scaler = 1.0f / ntfft;
  for(int iw = 0; iw < nw1; iw++) {
    const float f = (w_pow == 0) ? 1 : (iw == 0) ? 0 : powf(iw * dwf, w_pow);
    const std::complex<float> c = c_phase * f * scaler;
    fac[iw] = make_float2(c.real(), c.imag());
  }
  // traces per batch: the work space stays near 256 MB
  const size_t per = sizeof(float) * ((size_t)ntfft + 2 * (size_t)P->nw + 2 * (size_t)nw1);
  P->chunk = (int)std::max<size_t>(1, std::min<size_t>(65536, (256u << 20) / per));
  int rc = KLRTM_OK;
  if(hipMalloc(&P->d_tab, sizeof(float) * tab.size()) != hipSuccess || hipMalloc(&P->d_fac, sizeof(float2) * nw1) != hipSuccess ||
     hipMalloc(&P->d_pad, sizeof(float) * (size_t)P->chunk * ntfft) != hipSuccess ||
     hipMalloc(&P->d_spec, sizeof(float2) * (size_t)P->chunk * P->nw) != hipSuccess ||
     hipMalloc(&P->d_a, sizeof(float) * (size_t)P->chunk * 2 * nw1) != hipSuccess) {
    set_error("phase-correction work space does not fit");
    rc = KLRTM_ENOMEM;
  } else if(hipMemcpy(P->d_tab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(P->d_fac, fac.data(), sizeof(float2) * nw1, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("phase-correction table upload failed");
    rc = KLRTM_EDEVICE;
  } else if(rocblas_create_handle(&P->blas) != rocblas_status_success) {
    P->blas = nullptr;
    set_error("rocblas_create_handle failed");
    rc = KLRTM_EDEVICE;
  }
  if(rc) {
    delete P;
    return rc;
  }
  *out = P;
  return KLRTM_OK;
}

// applyBackward(trace, ntr, trcOut, scaler) on device traces [ntr][nt] -> [ntr][nt_dft]; the scaler is
// applied only when resampling, as the reference's loop over ntOut does
int phasecorr_backward(PhaseCorr &P, const float *in, int ntr, float *out, float scaler) {
  KL_REQUIRE(in && out && ntr >= 0, KLRTM_EINVAL, "bad traces");
  LaunchTimer tm("phase_correction");
  int rc;
  for(int t0 = 0; t0 < ntr; t0 += P.chunk) {
    const int m = std::min(P.chunk, ntr - t0);
    if(!P.have_plan || m != P.chunk) {
      if(P.have_plan) (void)hipfftDestroy(P.plan), P.have_plan = false;
      int n[1] = {P.ntfft};
      if(hipfftPlanMany(&P.plan, 1, n, nullptr, 1, P.ntfft, nullptr, 1, P.nw, HIPFFT_R2C, m) != HIPFFT_SUCCESS) {
        set_error("hipfftPlanMany (%d traces of %d) failed", m, P.ntfft);
        return KLRTM_EDEVICE;
      }
      P.have_plan = true;
    }
    hipLaunchKernelGGL(pc_pad_kernel, dim3(nblk((size_t)m * P.ntfft)), dim3(256), 0, stream(), in + (size_t)t0 * P.nt, P.d_pad,
                       P.nt, P.ntfft, m);
    if((rc = check_launch("pc_pad_kernel"))) return rc;
    if(hipfftSetStream(P.plan, stream()) != HIPFFT_SUCCESS ||
       hipfftExecR2C(P.plan, P.d_pad, (hipfftComplex *)P.d_spec) != HIPFFT_SUCCESS) {
      set_error("hipfftExecR2C failed");
      return KLRTM_EDEVICE;
    }
    hipLaunchKernelGGL(pc_spectrum_kernel, dim3(nblk((size_t)m * P.nw1)), dim3(256), 0, stream(), P.d_spec, P.d_fac, P.d_a,
                       P.nw, P.nw1, m);
    if((rc = check_launch("pc_spectrum_kernel"))) return rc;
    // out^T (nt_dft x m) = tab^T (nt_dft x 2nw1) * A^T (2nw1 x m), column-major views of the row-major arrays
    const float one = 1.0f, zero = 0.0f;
    float *o = out + (size_t)t0 * P.nt_dft;
    if(rocblas_set_stream(P.blas, stream()) != rocblas_status_success ||
       rocblas_sgemm(P.blas, rocblas_operation_none, rocblas_operation_none, P.nt_dft, m, 2 * P.nw1, &one, P.d_tab,
                     P.nt_dft, P.d_a, 2 * P.nw1, &zero, o, P.nt_dft) != rocblas_status_success) {
      set_error("rocblas_sgemm failed");
      return KLRTM_EDEVICE;
    }
    if(P.resample && scaler != 1.0f) {
      hipLaunchKernelGGL(pc_scale_kernel, dim3(nblk((size_t)m * P.nt_dft)), dim3(256), 0, stream(), o,
                         (size_t)m * P.nt_dft, scaler);
      if((rc = check_launch("pc_scale_kernel"))) return rc;
    }
  }
  return KLRTM_OK;
}

}  // namespace klrtm

using namespace klrtm;

extern "C" {
int klrtm_phasecorr_create(klrtm_phasecorr **p, int nt, float dt, float Dt, float t0, float fmax, int ntfft, float w_pow,
                           float phase_deg, int do_tdsp, int ntOut, float dtOut, float t0Out) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null output");
  PhaseCorr *P = nullptr;
  int rc = phasecorr_create(&P, nt, dt, Dt, t0, fmax, ntfft, w_pow, phase_deg, do_tdsp, ntOut, dtOut, t0Out);
  if(rc == KLRTM_OK) *p = reinterpret_cast<klrtm_phasecorr *>(P);
  return rc;
}
int klrtm_phasecorr_destroy(klrtm_phasecorr *p) {
  delete reinterpret_cast<PhaseCorr *>(p);
  return KLRTM_OK;
}
int klrtm_phasecorr_info(klrtm_phasecorr *p, int *nt_out, int *nw1) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null phase correction");
  const PhaseCorr &P = *reinterpret_cast<PhaseCorr *>(p);
  if(nt_out) *nt_out = P.nt_dft;
  if(nw1) *nw1 = P.nw1;
  return KLRTM_OK;
}
int klrtm_phasecorr_backward(klrtm_phasecorr *p, const float *in, int ntr, float *out, float scaler) {
  KL_REQUIRE(p, KLRTM_EINVAL, "null phase correction");
  return phasecorr_backward(*reinterpret_cast<PhaseCorr *>(p), in, ntr, out, scaler);
}
}
