// ImagingCorrelator.h -- bgprtm's ImagingCorrelator (bgprtm/ImagingCorrelator.h:17-60,
// bgprtm/ImagingCorrelator.cpp) over device volumes.
//
// The reference constructor ImagingCorrelator(nx,ny,nz,dx,dy,dz,nx2,ny2,nz2,kmin,nThreads) and
// run(fftv, image, imgFWI, imgRAW, srcWave, recWave, option) with the IMAGE_OPTION values
// (IMAGE1, INVERSION, INVERSION2, ALLIN3) run on the GPU at nDbl = 1 (nx2 = nx, ny2 = ny,
// nz2 = nz): the libFFTV interpolation is then the identity and `fftv` is not used (pass
// nullptr).  srcWave[0] / recWave[0] are device volumes the call may overwrite, as the
// reference's does.  interleave is a no-op at nDbl = 1 (nloop = 1).
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
class FFTVFilter;  // the reference's libFFTV filter (unvendored); unused at nDbl = 1
}

class ImagingCorrelator {
public:
  // bgprtm/ImagingCorrelator.cpp:17-83
  ImagingCorrelator(int nx, int ny, int nz, float dx, float dy, float dz, int nx2, int ny2, int nz2, float kmin,
                    int /*nThreads*/ = 0)
      : nx_(nx), ny_(ny), nz_(nz) {
    if(nx2 != nx || ny2 != ny || nz2 != nz)
      throw konglong::Error(KLRTM_EINVAL, "konglong: ImagingCorrelator runs at nDbl = 1 (nx2 = nx, ny2 = ny, nz2 = nz)");
    konglong::check(klrtm_correlator_create(&c_, nx, ny, nz, dx, dy, dz, kmin));
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
    konglong::check(klrtm_correlator_run(c_, image, imgFWI, imgRAW, nullptr, srcWave[0], recWave[0], option));
  }
  // nloop = 1 at nDbl = 1: nothing to interleave (bgprtm/ImagingCorrelator.cpp:441-446)
  void interleave(float * /*image*/, float ** /*srcBuf*/, float ** /*recBuf*/) {}
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
  int nx_, ny_, nz_;
  klrtm_correlator *c_ = nullptr;
  std::vector<float> wt_;
  const float *tz_ = nullptr;
  int ntz_ = 0;
};

#endif  // KONGLONG_IMAGINGCORRELATOR_H_
