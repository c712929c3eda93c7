// ImagingCorrelator.h -- the zero-lag imaging condition of bgprtm
// (ImagingCorrelator::crossone, bgprtm/ImagingCorrelator.cpp:180-194, with the
// source-side tapers and illumination of ImagingCondition::genImage/calIllum,
// bgprtm/ImagingCondition.cpp:143-226).  Device pointers.
//
// correlate(src, rec, it): image += S*R and illum += S*S with S = src times the
// time taper wtaperT[it] (it < ntaperT) or the water-depth z taper, as
// ImagingCondition::applyTape does.  FP64 image/illumination (for multi-shot
// stacks) accumulate the same products in double.
#ifndef KONGLONG_IMAGINGCORRELATOR_H_
#define KONGLONG_IMAGINGCORRELATOR_H_

#include <vector>

#include "konglong_gpu.h"

class ImagingCorrelator {
public:
  ImagingCorrelator(int nx, int ny, int nz) : nx_(nx), ny_(ny), nz_(nz) {}

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
  int nx_, ny_, nz_;
  std::vector<float> wt_;
  const float *tz_ = nullptr;
  int ntz_ = 0;
};

#endif  // KONGLONG_IMAGINGCORRELATOR_H_
