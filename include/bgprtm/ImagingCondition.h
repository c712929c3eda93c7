// ImagingCondition.h -- the image post-processing steps of bgprtm's
// ImagingCondition and its Lagrange resampler, over device volumes:
//   computeDamping / applyDamping (bgprtm/ImagingCondition.cpp:600-672)
//   imageConeMute (bgprtm/ImagingCondition.cpp:690-756)
//   Lagrange::apply (bgprtm/Lagrange.cpp:19-266)
//   saveImage onto an inline/crossline grid (bgprtm/ImagingCondition.cpp:874-945)
// and the free-surface mirror Propagator::apply_symmetry
// (bgprtm/Propagator.cpp:1857-1918).  global_pars keys become arguments with
// the reference's defaults.
#ifndef KONGLONG_IMAGINGCONDITION_H_
#define KONGLONG_IMAGINGCONDITION_H_

#include <vector>

#include "konglong_gpu.h"

namespace ImagingPost {

// ImagingCondition::computeDamping: illum becomes the damping weights in place; returns the average
inline double computeDamping(float *illum, int nx, int ny, int nz, bool smoothSrcAmp = true, int smoothSrcAmpXY = 10,
                             int smoothSrcAmpZ = 10, float ampDamping = 1.0f, float dampWeight = 0.07f) {
  double avg = 0.0;
  konglong::check(klrtm_illum_damping(illum, nx, ny, nz, smoothSrcAmp ? 1 : 0, smoothSrcAmpXY, smoothSrcAmpZ, ampDamping,
                                      dampWeight, &avg));
  return avg;
}
// ImagingCondition::applyDamping: img *= illum
inline void applyDamping(const float *illum, float *img, int nx, int ny, int nz) {
  konglong::check(klrtm_apply_damping(illum, img, (size_t)nx * ny * nz));
}

// ImagingCondition::imageConeMute on a rectangular grid; map: (depth, offset) pairs (empty: angle mute)
inline void imageConeMute(float *image, int nx, int ny, int nz, float x0, float y0, float z0, float dx, float dy,
                          float dz, float sourceX, float sourceY, float sourceZ, float angle = 30.0f,
                          float offset_min = 0.0f, const std::vector<float> &map = {}, float *weight = nullptr) {
  konglong::check(klrtm_cone_mute(image, weight, nx, ny, nz, x0, y0, z0, dx, dy, dz, sourceX, sourceY, sourceZ, angle,
                                  offset_min, map.empty() ? nullptr : map.data(), (int)map.size() / 2));
}

// ImagingCondition::saveImage onto a destination grid (do_stack: add, else copy into the overlap window)
inline void saveImage(const float *regimg, int nz, int nx, int ny, int ix0s, int iy0s, int incx, int incy, float *dest,
                      int ix0d, int iy0d, int nxd, int nyd, bool do_stack = true) {
  konglong::check(klrtm_save_image(regimg, nz, nx, ny, ix0s, iy0s, incx, incy, dest, ix0d, iy0d, nxd, nyd,
                                   do_stack ? 1 : 0));
}

// Propagator::apply_symmetry(wf, sym) with the surface at iz0 and the taper over the z boundary
inline void apply_symmetry(float *wf, int nx, int ny, int nz, int iz0, int sym, int tprlen, bool taper = true) {
  konglong::check(klrtm_apply_symmetry(wf, nx, ny, nz, iz0, sym, tprlen, taper ? 1 : 0));
}

}  // namespace ImagingPost

// a rectangular grid as the resampler needs it (bgprtm/Grid.h: nz, nx, ny, origins and spacings)
struct RectGrid {
  int nz, nx, ny;
  float z0, dz, x0, dx, y0, dy;
};

// Lagrange(gridIn, gridOut, nPoints).apply(dataIn, dataOut) on device volumes
class Lagrange {
public:
  Lagrange(const RectGrid &in, const RectGrid &out, int nPoints = 8) : in_(in), out_(out), np_(nPoints) {}
  void apply(const float *dataIn, float *dataOut) const {
    konglong::check(klrtm_lagrange_resample(dataIn, in_.nz, in_.nx, in_.ny, in_.z0, in_.dz, in_.x0, in_.dx, in_.y0,
                                            in_.dy, dataOut, out_.nz, out_.nx, out_.ny, out_.z0, out_.dz, out_.x0,
                                            out_.dx, out_.y0, out_.dy, np_));
  }

private:
  RectGrid in_, out_;
  int np_;
};

#endif  // KONGLONG_IMAGINGCONDITION_H_
