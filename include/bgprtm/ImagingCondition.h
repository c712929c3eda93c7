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

#include <cmath>
#include <memory>
#include <vector>

#include "CacheFile.h"
#include "ImagingCorrelator.h"
#include "Propagator.h"
#include "WaveFieldCompress.h"
#include "konglong_gpu.h"

// bgprtm/ImagingCondition.h:32-34
#ifndef KONGLONG_NO_IMAGETYPE
enum ImageType { STACK = 0, OFCIG = 1, DTCIG = 2, ADCIG = 3, DECON = 4 };
#endif

namespace konglong {
// global_pars keys ImagingCondition reads (bgprtm/ImagingCondition.cpp:38-70, 151-198, 393-396): the
// imaging option (LowcutFilter), the water-depth taper length (grid->getIDz(waterDepth)), minFreq / vMin
// and the imaging grid doubling (ic_double_z/x/y with highTapeZ / highTapeXY, :61-67, 169-178; the reference's
// default doubles when the FD dispersion exceeds 0.6, here the default is 1)
struct ImagingParams {
  int LowcutFilter = KLRTM_IC_IMAGE1;
  int ntaperZ = 0;
  float minFreq = 1.0f, vMin = 1492.f;
  int ic_double_z = 1, ic_double_x = 1, ic_double_y = 1;
  float highTapeZ = 0.9f, highTapeXY = 0.7f;
};
inline ImagingParams &imaging_params() {
  static ImagingParams p;
  return p;
}
}  // namespace konglong

// ImagingCondition(srcWaveField, recWaveField, imgType)::genImage (bgprtm/ImagingCondition.cpp:322-470) over two
// snapshot caches: for every stored step the source and receiver snapshots are read from their caches,
// decoded with WaveFieldCompress::compression (with its edge taper), the source one tapered in time
// (wtaperT over ntaperT = rint((t1 - t0)/dt) steps) or by the water-depth taper, the illumination (AMP)
// accumulated and ImagingCorrelator::run applied with the LowcutFilter option; INVERSION2 then takes its
// cos and sin terms.  Zero lag only (MaxTimeLag 0, STACK).  On a doubled imaging grid (IMAGE1 or INVERSION) each
// step high-cuts S and R (applyFFTV), accumulates the high-cut S^2, forms the half-sample copies and correlates
// them into sub-images (stackImageInterp, :104-141), which are interleaved at the end, the illumination
// averaged onto the doubled grid (:415-446).  The image (and AMP, FWI, RAW) come back on the host in the
// returned Grid (the doubled grid's sizes and spacings) and the accessors; with destImages and an OutputGrid the image is
// added into destImages[0] where its columns fall on output nodes.  Cone mute, illumination damping and
// resampling are the separate ImagingPost / Lagrange calls.
class ImagingCondition {
public:
  ImagingCondition(CacheFile *srcWaveField, CacheFile *recWaveField, ImageType imgType = STACK)
      : src_(srcWaveField), rec_(recWaveField) {
    if(imgType != STACK) throw konglong::Error(KLRTM_EINVAL, "konglong: only the zero-lag STACK image is built");
    if(!src_ || !rec_ || src_->nx != rec_->nx || src_->ny != rec_->ny || src_->nz != rec_->nz || src_->nt != rec_->nt)
      throw konglong::Error(KLRTM_EINVAL, "konglong: source and receiver caches must have the same shape");
  }
  virtual ~ImagingCondition() = default;
  void getVel(float *vel, float /*dt*/) { vel_ = vel; }

  std::shared_ptr<Grid> genImage(Grid *grid, WaveFieldCompress *wfComp, float /*xMinValid*/, float /*xMaxValid*/,
                                 float /*yMinValid*/, float /*yMaxValid*/, float /*dtsav*/,
                                 std::vector<float *> destImages = std::vector<float *>{}, OutputGrid *destGrid = nullptr) {
    const int nx = src_->nx, ny = src_->ny, nz = src_->nz, nt = src_->nt;
    const size_t n = (size_t)nx * ny * nz, ns = WaveFieldCompress::nshort_volume(nz, nx, ny);
    const konglong::ImagingParams &P = konglong::imaging_params();
    const int opt = P.LowcutFilter;
    const int dz = P.ic_double_z, dx = nx > 1 ? P.ic_double_x : 1, dy = ny > 1 ? P.ic_double_y : 1;
    if(dz * dx * dy > 1) return genImageDouble(grid, wfComp, dz, dx, dy, destImages, destGrid);
    // ImagingCondition::calTapes (bgprtm/ImagingCondition.cpp:180-198)
    std::vector<float> wtaperT;
    const int ntaperT = (int)nearbyintf((src_->t1 - src_->t0) / src_->dt);
    for(int i = 0; i < ntaperT; i++) wtaperT.push_back(0.5f * (1.0f - cosf((float)M_PI * float(i) / float(ntaperT))));
    std::vector<float> wtaperZ(std::max(P.ntaperZ, 0));
    for(int i = 0; i < P.ntaperZ; i++) wtaperZ[i] = 0.5f * (1.0f - cosf((float)M_PI * float(i) / float(P.ntaperZ)));
    konglong::DeviceBuffer<float> dtz(std::max<size_t>(1, wtaperZ.size()));
    if(!wtaperZ.empty()) dtz.upload(wtaperZ.data());
    konglong::DeviceBuffer<ushort> cs(ns), cr(ns);
    konglong::DeviceBuffer<float> S(n), R(n), img(n), amp(n), fwi(n), raw(n);
    std::vector<ushort> hs(ns), hr(ns);
    std::unique_ptr<ImagingCorrelator> corr;
    if(opt != KLRTM_IC_IMAGE1)
      corr.reset(new ImagingCorrelator(nx, ny, nz, src_->dx, src_->dy, src_->dz, nx, ny, nz, P.minFreq / P.vMin));
    for(int it = 0; it < nt; it++) {
      const off_t off = (off_t)it * (off_t)(ns * sizeof(ushort));
      if(src_->pread(hs.data(), ns * sizeof(ushort), off) < 0 || rec_->pread(hr.data(), ns * sizeof(ushort), off) < 0)
        throw konglong::Error(KLRTM_EINVAL, "konglong: short read from a snapshot cache");
      cs.upload(hs.data()), cr.upload(hr.data());
      wfComp->uncompress(cs.get(), S.get(), nx, ny, nz);
      wfComp->uncompress(cr.get(), R.get(), nx, ny, nz);
      // applyTape, then calIllum (amp += S*S) and ImagingCorrelator::run
      konglong::check(klrtm_apply_tape(S.get(), nx, ny, nz, it < ntaperT ? wtaperT[it] : -1.0f, dtz.get(),
                                       it < ntaperT ? 0 : P.ntaperZ));
      konglong::check(klrtm_ic_correlate(S.get(), S.get(), nx, ny, nz, 1.0f, nullptr, 0, amp.get(), nullptr, nullptr,
                                         nullptr));
      if(opt == KLRTM_IC_IMAGE1) {
        konglong::check(klrtm_ic_correlate(S.get(), R.get(), nx, ny, nz, 1.0f, nullptr, 0, img.get(), nullptr, nullptr,
                                           nullptr));
      } else {
        float *sb[1] = {S.get()}, *rb[1] = {R.get()};
        corr->run(nullptr, img.get(), fwi.get(), raw.get(), sb, rb, opt);
      }
    }
    if(opt == KLRTM_IC_INVERSION2) corr->costerm(img.get()), corr->sinterm(fwi.get(), img.get());
    image_ = img.host(), amp_ = amp.host(), fwi_ = fwi.host(), raw_ = raw.host();
    return finish(grid, 1, 1, 1, destImages, destGrid);
  }
  const std::vector<float> &image() const { return image_; }
  const std::vector<float> &illum() const { return amp_; }
  const std::vector<float> &fwi() const { return fwi_; }
  const std::vector<float> &raw() const { return raw_; }

private:
  std::shared_ptr<Grid> genImageDouble(Grid *grid, WaveFieldCompress *wfComp, int dz, int dx, int dy,
                                       std::vector<float *> &destImages, OutputGrid *destGrid) {
    const int nx = src_->nx, ny = src_->ny, nz = src_->nz, nt = src_->nt;
    const size_t n = (size_t)nx * ny * nz, ns = WaveFieldCompress::nshort_volume(nz, nx, ny);
    const konglong::ImagingParams &P = konglong::imaging_params();
    const int opt = P.LowcutFilter;
    if(opt != KLRTM_IC_IMAGE1 && opt != KLRTM_IC_INVERSION)
      throw konglong::Error(KLRTM_EINVAL, "konglong: the doubled imaging grid takes IMAGE1 and INVERSION");
    ImagingCorrelator corr(nx, ny, nz, src_->dx, src_->dy, src_->dz, nx * dx, ny * dy, nz * dz, P.minFreq / P.vMin);
    konglong::check(klrtm_correlator_set_double(corr.handle(), dz, dx, dy, P.highTapeZ, P.highTapeXY));
    const int nl = corr.nloop();
    std::vector<float> wtaperT;
    const int ntaperT = (int)nearbyintf((src_->t1 - src_->t0) / src_->dt);
    for(int i = 0; i < ntaperT; i++) wtaperT.push_back(0.5f * (1.0f - cosf((float)M_PI * float(i) / float(ntaperT))));
    std::vector<float> wtaperZ(std::max(P.ntaperZ, 0));
    for(int i = 0; i < P.ntaperZ; i++) wtaperZ[i] = 0.5f * (1.0f - cosf((float)M_PI * float(i) / float(P.ntaperZ)));
    konglong::DeviceBuffer<float> dtz(std::max<size_t>(1, wtaperZ.size()));
    if(!wtaperZ.empty()) dtz.upload(wtaperZ.data());
    konglong::DeviceBuffer<ushort> cs(ns), cr(ns);
    konglong::DeviceBuffer<float> S(n * nl), R(n * nl), img(n * nl), amp(n), fwi(n * nl), amp2(n * nl);
    std::vector<ushort> hs(ns), hr(ns);
    std::vector<float *> sb(nl), rb(nl);
    for(int i = 0; i < nl; i++) sb[i] = S.get() + i * n, rb[i] = R.get() + i * n;
    for(int it = 0; it < nt; it++) {
      const off_t off = (off_t)it * (off_t)(ns * sizeof(ushort));
      if(src_->pread(hs.data(), ns * sizeof(ushort), off) < 0 || rec_->pread(hr.data(), ns * sizeof(ushort), off) < 0)
        throw konglong::Error(KLRTM_EINVAL, "konglong: short read from a snapshot cache");
      cs.upload(hs.data()), cr.upload(hr.data());
      wfComp->uncompress(cs.get(), S.get(), nx, ny, nz);
      wfComp->uncompress(cr.get(), R.get(), nx, ny, nz);
      konglong::check(klrtm_apply_tape(S.get(), nx, ny, nz, it < ntaperT ? wtaperT[it] : -1.0f, dtz.get(),
                                       it < ntaperT ? 0 : P.ntaperZ));
      // applyFFTV, calIllum on the high-cut S, interp (IMAGE1) and run
      konglong::check(klrtm_correlator_hightape(corr.handle(), S.get()));
      konglong::check(klrtm_correlator_hightape(corr.handle(), R.get()));
      konglong::check(klrtm_ic_correlate(S.get(), S.get(), nx, ny, nz, 1.0f, nullptr, 0, amp.get(), nullptr, nullptr,
                                         nullptr));
      if(opt == KLRTM_IC_IMAGE1) {
        konglong::check(klrtm_correlator_interp(corr.handle(), S.get()));
        konglong::check(klrtm_correlator_interp(corr.handle(), R.get()));
      }
      corr.run(nullptr, img.get(), fwi.get(), nullptr, sb.data(), rb.data(), opt);
    }
    corr.interleave(img.get(), sb.data(), rb.data());
    corr.interleave(fwi.get(), sb.data(), rb.data());
    konglong::check(klrtm_correlator_interp_illum(corr.handle(), amp.get(), amp2.get()));
    image_ = img.host(), amp_ = amp2.host(), fwi_ = fwi.host(), raw_.assign(n * nl, 0.f);
    return finish(grid, dz, dx, dy, destImages, destGrid);
  }

  // the returned grid, and the image added into destImages[0] on the output nodes
  std::shared_ptr<Grid> finish(Grid *grid, int dz, int dx, int dy, std::vector<float *> &destImages,
                               OutputGrid *destGrid) {
    const int nx = src_->nx * dx, ny = src_->ny * dy, nz = src_->nz * dz;
    auto out = std::make_shared<Grid>();
    if(grid) *out = *grid;
    out->nx = nx, out->ny = ny, out->nz = nz, out->image = image_;
    out->dx /= dx, out->dy /= dy, out->dz /= dz;
    if(destGrid && !destImages.empty() && destImages[0]) {
      // columns on output nodes (the grid's origin and spacing place them), depth sampling shared
      for(int iy = 0; iy < ny; iy++)
        for(int ix = 0; ix < nx; ix++) {
          const double fx = (out->x0 + ix * out->dx - destGrid->x0) / destGrid->dx;
          const double fy = destGrid->ny > 1 ? (out->y0 + iy * out->dy - destGrid->y0) / destGrid->dy : 0.0;
          const long jx = std::lround(fx), jy = std::lround(fy);
          if(std::fabs(fx - jx) > 1e-3 || std::fabs(fy - jy) > 1e-3 || jx < 0 || jy < 0 || jx >= destGrid->nx ||
             jy >= destGrid->ny)
            continue;
          float *col = destImages[0] + ((size_t)jy * destGrid->nx + jx) * destGrid->nz;
          const float *srcc = image_.data() + ((size_t)iy * nx + ix) * nz;
          for(int iz = 0; iz < std::min(nz, destGrid->nz); iz++) col[iz] += srcc[iz];
        }
    }
    return out;
  }

  CacheFile *src_, *rec_;
  float *vel_ = nullptr;
  std::vector<float> image_, amp_, fwi_, raw_;
};

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
