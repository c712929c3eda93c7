// ImagingCorrelator.h -- bgprtm's ImagingCorrelator (bgprtm/ImagingCorrelator.h:17-60,
// bgprtm/ImagingCorrelator.cpp) over device volumes.
//
// The reference constructor ImagingCorrelator(nx,ny,nz,dx,dy,dz,nx2,ny2,nz2,kmin,nThreads) and
// run(fftv, image, imgFWI, imgRAW, srcWave, recWave, option) with the IMAGE_OPTION values
// (IMAGE1, INVERSION, INVERSION2, ALLIN3) run on the GPU; `fftv` is not used (pass nullptr): the
// half-sample interpolation is this library's.  srcWave[i] / recWave[i] are device volumes the call
// may overwrite, as the reference's does.  On a doubled grid (nx2 = 2 nx, ny2 = 2 ny, nz2 = 2 nz in
// 3D; nz2 = 2 nz and nx2 = nx or 2 nx in 2D) the options are IMAGE1 (srcWave[i] the copies the caller
// formed with ImagingCondition::interp) and INVERSION (srcWave[0] given, the copies formed here), the
// nloop() copies must lie back to back (srcWave[i] = srcWave[0] + i*nx*ny*nz), the images hold nloop()
// sub-images, and interleave() puts them onto the doubled grid in place, using srcBuf as scratch as
// the reference does (bgprtm/ImagingCorrelator.cpp:440-513).
//
// The zero-lag helpers of this build stay: correlate(src, rec, it) applies the source tapers of
// ImagingCondition::applyTape and accumulates illumination (ImagingCondition::calIllum) in the
// same pass; crossone is ImagingCorrelator::crossone with isign = +1.
#ifndef KONGLONG_IMAGINGCORRELATOR_H_
#define KONGLONG_IMAGINGCORRELATOR_H_

#include <vector>

#include "konglong_gpu.h"

// bgprtm/ImagingCorrelator.h:17-19
#ifndef KONGLONG_NO_IMAGE_OPTION
enum IMAGE_OPTION { IMAGE1 = KLRTM_IC_IMAGE1, INVERSION = KLRTM_IC_INVERSION, INVERSION2 = KLRTM_IC_INVERSION2,
                    ALLIN3 = KLRTM_IC_ALLIN3 };
#endif

namespace libfftv {
class FFTVFilter;  // the reference's libFFTV filter (unvendored); not used
}

class ImagingCorrelator {
public:
  // bgprtm/ImagingCorrelator.cpp:17-83
  ImagingCorrelator(int nx, int ny, int nz, float dx, float dy, float dz, int nx2, int ny2, int nz2, float kmin,
                    int /*nThreads*/ = 0)
      : nx_(nx), ny_(ny), nz_(nz) {
    konglong::check(klrtm_correlator_create(&c_, nx, ny, nz, dx, dy, dz, kmin));
    if(nx2 != nx || ny2 != ny || nz2 != nz) {
      if(nx2 % nx || ny2 % ny || nz2 % nz) {
        klrtm_correlator_destroy(c_), c_ = nullptr;
        throw konglong::Error(KLRTM_EINVAL, "konglong: nx2, ny2, nz2 must be 1 or 2 times nx, ny, nz");
      }
      const int rc = klrtm_correlator_set_double(c_, nz2 / nz, nx2 / nx, ny2 / ny, 0.9f, 0.7f);
      if(rc) {
        klrtm_correlator_destroy(c_), c_ = nullptr;
        konglong::check(rc);
      }
    }
    konglong::check(klrtm_correlator_nloop(c_, &nloop_));
  }
  // the zero-lag correlator of this build (no FFT work space)
  ImagingCorrelator(int nx, int ny, int nz) : nx_(nx), ny_(ny), nz_(nz) {}
  virtual ~ImagingCorrelator() {
    if(c_) klrtm_correlator_destroy(c_);
  }
  ImagingCorrelator(const ImagingCorrelator &) = delete;
  ImagingCorrelator &operator=(const ImagingCorrelator &) = delete;

  // bgprtm/ImagingCorrelator.cpp:355-364
  void run(libfftv::FFTVFilter * /*fftv*/, float *image, float *imgFWI, float *imgRAW, float **srcWave, float **recWave,
           int option) {
    need_fft();
    if(nloop_ == 1) {
      konglong::check(klrtm_correlator_run(c_, image, imgFWI, imgRAW, nullptr, srcWave[0], recWave[0], option));
      return;
    }
    need_back_to_back(srcWave), need_back_to_back(recWave);
    konglong::check(klrtm_correlator_run_copies(c_, image, imgFWI, srcWave[0], recWave[0], option));
  }
  // bgprtm/ImagingCorrelator.cpp:440-445: image (nloop sub-images) onto the doubled grid in place
  void interleave(float *image, float **srcBuf, float ** /*recBuf*/) {
    if(!image || nloop_ == 1) return;
    need_back_to_back(srcBuf);
    const size_t bytes = sizeof(float) * nloop_ * (size_t)nx_ * ny_ * nz_;
    konglong::check(klrtm_memcpy_d2d(srcBuf[0], image, bytes));
    konglong::check(klrtm_correlator_interleave(c_, srcBuf[0], image));
  }
  int nloop() const { return nloop_; }
  klrtm_correlator *handle() const { return c_; }
  // bgprtm/ImagingCorrelator.cpp:308-342
  void costerm(float *image) {
    need_fft();
    konglong::check(klrtm_correlator_costerm(c_, image));
  }
  void sinterm(float *imgfull, float *imgcos) {
    need_fft();
    konglong::check(klrtm_correlator_sinterm(c_, imgfull, imgcos));
  }

  // wtaperT: time taper (ntaperT steps); wtaperZ: water-depth taper over the first ntaperZ cells (device)
  void setTapers(const std::vector<float> &wtaperT, const float *wtaperZ_device, int ntaperZ) {
    wt_ = wtaperT, tz_ = wtaperZ_device, ntz_ = ntaperZ;
  }
  // zero-lag correlation of one stored step; any output may be null
  void correlate(const float *src, const float *rec, int it, float *image, float *illum, double *image64 = nullptr,
                 double *illum64 = nullptr) {
    // taper_t >= 0 selects the time taper, < 0 the water-depth taper
    const bool time_taper = it < (int)wt_.size();
    const float t = time_taper ? wt_[it] : -1.0f;
    konglong::check(klrtm_ic_correlate(src, rec, nx_, ny_, nz_, t, time_taper ? nullptr : tz_,
                                       time_taper ? 0 : ntz_, image, illum, image64, illum64));
  }
  // ImagingCorrelator::crossone: image += src*rec (no tapers)
  void crossone(float *image, const float *src, const float *rec) {
    konglong::check(klrtm_ic_correlate(src, rec, nx_, ny_, nz_, -1.0f, nullptr, 0, image, nullptr, nullptr, nullptr));
  }

private:
  void need_fft() const {
    if(!c_) throw konglong::Error(KLRTM_ESTATE, "konglong: the FFT imaging conditions need the reference constructor");
  }
  void need_back_to_back(float **buf) const {
    for(int i = 1; i < nloop_; i++)
      if(buf[i] != buf[0] + (size_t)i * nx_ * ny_ * nz_)
        throw konglong::Error(KLRTM_EINVAL, "konglong: the doubled-grid copies must lie back to back");
  }
  int nx_, ny_, nz_;
  int nloop_ = 1;
  klrtm_correlator *c_ = nullptr;
  std::vector<float> wt_;
  const float *tz_ = nullptr;
  int ntz_ = 0;
};

#endif  // KONGLONG_IMAGINGCORRELATOR_H_
