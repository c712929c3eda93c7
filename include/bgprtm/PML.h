// PML.h -- bgprtm's PML class (bgprtm/PML.h:28-61) with device-side applies.
//
// The constructor builds coef/feoc on the host exactly as PML::init
// (bgprtm/PML.cpp:23-60) and keeps device copies; the apply methods take
// device pointers and run the reference rounds (bit-identical to the FP32
// CPU code).  pml_eps comes from konglong::params().
#ifndef KONGLONG_PML_H_
#define KONGLONG_PML_H_

#include <vector>

#include "konglong_gpu.h"

class PML {
public:
  PML(int n, int npml, int nbatch, float dx, float dt, float vmax, bool limit_slope)
      : n(n), npml(npml), nbatch(nbatch), dx(dx), dt(dt), vmax(vmax), limit_slope(limit_slope), coef(npml),
        feoc(npml), dcoef_(npml), dfeoc_(npml) {
    if(limit_slope) throw konglong::Error(KLRTM_EINVAL, "konglong: limit_slope profiles are not supported");
    konglong::check(klrtm_pml_profile(npml, dx, dt, vmax, konglong::params().pml_eps, coef.data(), feoc.data()));
    dcoef_.upload(coef.data());
    dfeoc_.upload(feoc.data());
  }
  virtual ~PML() = default;

  // one line of npml samples with profile c (this->coef, this->feoc or another host vector) times
  // scaler; then p[npml] = p[npml-1]*symmetry when symmetry != 0
  void apply_single(std::vector<float> &c, float scaler, float *p, float *q, int symmetry = 0) {
    konglong::check(klrtm_pml_apply_lines(device_profile(c), npml, 1, n, scaler, p, q, nullptr, symmetry));
  }
  void apply2_single(std::vector<float> &c, float *p, float *q, float *q2, int symmetry = 0) {
    konglong::check(klrtm_pml_apply_lines(device_profile(c), npml, 1, n, 1.0f, p, q, q2, symmetry));
  }
  void apply_single_trans(float g, float *p, float *q, int nz) {
    konglong::check(klrtm_pml_apply_trans(g, p, q, nullptr, nz));
  }
  void apply2_single_trans(float g, float *p, float *q, float *q2, int nz) {
    konglong::check(klrtm_pml_apply_trans(g, p, q, q2, nz));
  }
  // topbot -1: top (coef, layer at the start of each line), 1: bottom (feoc), 2/3: bottom with the
  // (anti)symmetric sample mirrored one point early (bgprtm/PML.cpp:88-119)
  void apply(float *pwav, float *qpml, int nb, int topbot) { lines(pwav, qpml, nullptr, nb, topbot); }
  void apply2(float *pwav, float *qpml, float *q2pml, int nb, int topbot) { lines(pwav, qpml, q2pml, nb, topbot); }

  const float *coef_device() const { return dcoef_.get(); }
  const float *feoc_device() const { return dfeoc_.get(); }

  int n, npml, nbatch;
  float dx, dt, vmax;
  bool limit_slope;
  std::vector<float> coef, feoc;

private:
  const float *device_profile(std::vector<float> &c) {
    if(&c == &coef) return dcoef_.get();
    if(&c == &feoc) return dfeoc_.get();
    if((int)c.size() < npml) throw konglong::Error(KLRTM_EINVAL, "konglong: PML profile shorter than npml");
    if(dother_.size() < (size_t)npml) dother_ = konglong::DeviceBuffer<float>(npml);
    konglong::check(klrtm_memcpy_h2d(dother_.get(), c.data(), npml * sizeof(float)));
    return dother_.get();
  }
  void lines(float *p, float *q, float *q2, int nb, int topbot) {
    const int sym = (topbot == 2) ? 1 : (topbot == 3) ? -1 : 0;
    if(topbot == -1) {
      konglong::check(klrtm_pml_apply_lines(dcoef_.get(), npml, nb, n, 1.0f, p, q, q2, sym));
    } else if(topbot > 0) {
      float *pb = p + (n - npml) - ((topbot == 2 || topbot == 3) ? 1 : 0);
      konglong::check(klrtm_pml_apply_lines(dfeoc_.get(), npml, nb, n, 1.0f, pb, q, q2, sym));
    }
  }
  konglong::DeviceBuffer<float> dcoef_, dfeoc_, dother_;
};

#endif  // KONGLONG_PML_H_
