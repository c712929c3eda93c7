// Propagator.h -- one shot of RTM or modelling (Propagator::rtm / Propagator::mod,
// bgprtm/Propagator.h:100-104; the time loop of Propagator::migration,
// bgprtm/Propagator.cpp:1157-1260) on the GPU.
//
// The reference builds its configuration from global_pars and its model
// loaders; here a klrtm_shot_config carries the same quantities.  Volumes are
// host arrays in the reference layout (iy*nx + ix)*nz + iz.
#ifndef KONGLONG_PROPAGATOR_H_
#define KONGLONG_PROPAGATOR_H_

#include <vector>

#include "konglong_gpu.h"

class Propagator {
public:
  explicit Propagator(const klrtm_shot_config &cfg) : cfg_(cfg) { konglong::check(klrtm_propagator_create(&p_, &cfg)); }
  virtual ~Propagator() { klrtm_propagator_destroy(p_); }
  Propagator(const Propagator &) = delete;
  Propagator &operator=(const Propagator &) = delete;

  // velocity in m/s (the Vel2 transform happens inside), optional eps, del, pjx, pjy
  void setModel(const float *vel, const float *eps = nullptr, const float *del = nullptr, const float *pjx = nullptr,
                const float *pjy = nullptr) {
    konglong::check(klrtm_propagator_set_model(p_, vel, eps, del, pjx, pjy));
  }
  // Q attenuation (Model::useQ): 1/Q volume and Q::coeffs [cq0, cq[order], wq[order]]; call before
  // setModel.  An empty coefficient vector switches Q off.
  void setQ(const float *invQ, const std::vector<float> &coeffs) {
    const int order = coeffs.empty() ? 0 : (int)(coeffs.size() - 1) / 2;
    konglong::check(klrtm_propagator_set_q(p_, invQ, order, coeffs.empty() ? nullptr : coeffs.data()));
  }
  // density (Model::useRho): host volume, nullptr switches it off
  void setRho(const float *rho) { konglong::check(klrtm_propagator_set_rho(p_, rho)); }
  // source array (Propagator::setupReceiver): positions xyz[3*n] in grid coordinates, delays in seconds
  void setSourceArray(const std::vector<float> &xyz, const std::vector<float> &delay) {
    if(xyz.size() != 3 * delay.size()) throw konglong::Error(KLRTM_EINVAL, "konglong: xyz needs 3 values per delay");
    konglong::check(klrtm_propagator_set_source_array(p_, (int)delay.size(), xyz.data(), delay.data()));
  }
  // SOURCE_GHOST / RECEIVER_GHOST / GHOST: mirror the chosen sides' injection about iz_mirror
  void setGhost(int izMirror, bool sourceSide, bool receiverSide) {
    konglong::check(klrtm_propagator_set_ghost(p_, izMirror, sourceSide ? 1 : 0, receiverSide ? 1 : 0));
  }
  // CacheFile: snapshots to host memory (when they fit memLimitGB) or to the file at path
  void setSnapshotCache(const char *path, double memLimitGB, bool keep = false) {
    konglong::check(klrtm_propagator_set_snapshot_cache(p_, path, memLimitGB, keep ? 1 : 0));
  }
  void setReceivers(const std::vector<float> &xyz) { konglong::check(klrtm_propagator_set_receivers(p_, xyz.data())); }
  void setSourceWavelet(const std::vector<float> &w) {
    konglong::check(klrtm_propagator_set_source_wavelet(p_, w.data()));
  }
  // Propagator::mod: the shot record [nrcv][nt]
  std::vector<float> mod() {
    std::vector<float> rec((size_t)cfg_.nrcv * cfg_.nt);
    konglong::check(klrtm_propagator_model(p_, rec.data()));
    return rec;
  }
  // Propagator::rtm: image (and illumination) of the record [nrcv][nt]; FP64 outputs for stacking
  void rtm(const std::vector<float> &record, float *image, float *illum = nullptr, double *image64 = nullptr,
           double *illum64 = nullptr) {
    konglong::check(klrtm_propagator_rtm(p_, record.data(), image, illum, image64, illum64));
  }
  void timings(double &fwd_ms, double &bwd_ms, double &ic_ms) {
    konglong::check(klrtm_propagator_timings(p_, &fwd_ms, &bwd_ms, &ic_ms));
  }

private:
  klrtm_shot_config cfg_;
  klrtm_propagator *p_ = nullptr;
};

#endif  // KONGLONG_PROPAGATOR_H_
