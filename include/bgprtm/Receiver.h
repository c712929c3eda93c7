// Receiver.h -- source/receiver injection and extraction
// (Receiver::spreadCoeffs/apply, bgprtm/Receiver.cpp:306-507,1204-1279;
// ExtractReceiverData::extract, bgprtm/ExtractReceiverData.cpp:199-229).
//
// Coordinates are float grid indices (x, y, z); traces are [nr][nt] on the
// host; wavefields and vel are device pointers.
#ifndef KONGLONG_RECEIVER_H_
#define KONGLONG_RECEIVER_H_

#include <vector>

#include "konglong_gpu.h"

class Receiver {
public:
  // spread: kernel half length (taps 2*spread); k1 <= k2 in (0, 0.5]: the spread band (khpass, khcut)
  Receiver(int nx, int ny, int nz, const std::vector<float> &xyz, int spread, float k1, float k2, int nt, float dx,
           float dy, float dz)
      : nr_((int)(xyz.size() / 3)), nt_(nt) {
    const float dt_dxyz = ny > 1 ? 1.0f / (dz * dx * dy) : 1.0f / (dz * dx);  // bgprtm/Receiver.cpp:55-57
    konglong::check(klrtm_receiver_create(&r_, nx, ny, nz, nr_, xyz.data(), spread, k1, k2, nt, dt_dxyz));
  }
  virtual ~Receiver() { klrtm_receiver_destroy(r_); }
  Receiver(const Receiver &) = delete;
  Receiver &operator=(const Receiver &) = delete;

  void setTraces(const std::vector<float> &traces) { konglong::check(klrtm_receiver_set_traces(r_, traces.data())); }
  // w0 += trace[:, it] * scaler * dt_dxyz * kx*ky*kz * vel
  void apply(float *w0, int it, const float *vel, float scaler = 1.0f) {
    konglong::check(klrtm_receiver_apply(r_, w0, it, vel, scaler));
  }
  // Receiver::apply(wf, it, ghost, vdt2, scaler0) (bgprtm/Receiver.cpp:1204-1279): with ghost the
  // injection is mirrored about the surface plane.  The reference finds that plane itself from its grid
  // and zsurf (iz_mirror = grid->getIDz(zsurf), bgprtm/Receiver.cpp:1204); this build takes it as
  // izMirror, because it has no Grid object
  void apply(float *w0, int it, bool ghost, int izMirror, const float *vel, float scaler = 1.0f) {
    if(ghost) konglong::check(klrtm_receiver_apply_ghost(r_, w0, it, vel, scaler, izMirror));
    else konglong::check(klrtm_receiver_apply(r_, w0, it, vel, scaler));
  }
  // record[:, jt] = sum kx*ky*kz * w (minus the mirrored samples with a ghost plane izMirror >= 0): the
  // reference's per-receiver extract_value(wf, ir, mirror_ghost) (bgprtm/Receiver.cpp:394-426) for every
  // receiver at once, with the plane passed in as for apply
  void extract(const float *w, int jt, int izMirror = -1) {
    if(izMirror >= 0) konglong::check(klrtm_receiver_extract_ghost(r_, w, jt, izMirror));
    else konglong::check(klrtm_receiver_extract(r_, w, jt));
  }
  std::vector<float> record() const {
    std::vector<float> rec((size_t)nr_ * nt_);
    konglong::check(klrtm_receiver_get_record(r_, rec.data()));
    return rec;
  }
  int size() const { return nr_; }
  klrtm_receiver *handle() const { return r_; }

private:
  klrtm_receiver *r_ = nullptr;
  int nr_, nt_;
};

#endif  // KONGLONG_RECEIVER_H_
