// ExtractReceiverData.h -- the modelled shot record of bgprtm's ExtractReceiverData
// (bgprtm/ExtractReceiverData.h:19-80): the raw record the propagator extracts, its resampling with
// PhaseCorrection (resampleRecData, bgprtm/ExtractReceiverData.cpp:75-110) and the record file
// (saveRecFile, bgprtm/ExtractReceiverData.cpp:263-381).
//
// The reference constructor takes the modelling Grid and the boundary type; here the receivers and the
// extraction are the Receiver / Propagator classes (Propagator::mod extracts every step), and this class
// holds what follows: recDataRaw [nr][nt] at dt from t0 (t0 <= 0), resampled on the GPU to oData
// [nr][ntOut] at dtOut.  global_pars keys become arguments: outputDt / outputNt (setResampleNt),
// _dt_prop and maxFreq (resampleRecData), modeling_keep_t0.  The record file is this build's
// (klrtm_record_write), since the reference's jseisIO / libSWIO writers are not vendored.
#ifndef KONGLONG_EXTRACTRECEIVERDATA_H_
#define KONGLONG_EXTRACTRECEIVERDATA_H_

#include <cmath>
#include <vector>

#include "konglong_gpu.h"

class ExtractReceiverData {
public:
  // nr receivers, nt raw samples at dt starting at t0; tmax is the output length (global_pars tmax)
  ExtractReceiverData(int nr, int nt, float dt, float t0, float tmax)
      : nt(nt), ntOut((int)nearbyintf(tmax / dt) + 1), nr(nr), dt(dt), dtOut(dt), t0(t0), tmax(tmax),
        recDataRaw((size_t)nr * nt, 0.f) {}
  virtual ~ExtractReceiverData() = default;

  void setNt(int n, float t) {
    nt = n, t0 = t;
    recDataRaw.assign((size_t)nr * nt, 0.f);
  }
  void setResampleNt(float resampleDt, int resampleNt) {
    dtOut = resampleDt, ntOut = resampleNt, tmax = (resampleNt - 1) * resampleDt;
  }

  // resampleRecData(dat): PhaseCorrection(nt, dt, dt_prop, -t0, maxFreq, BWD, w_pow, phase_deg, 1, ntOut,
  // dtOut, keep_t0 ? -t0 : 0) per trace, scaled by 1/dt_spread; a dipole spread (did_dipole) uses
  // w^-1, 90 degrees and dt_spread = 0.5 dz / vsurf
  void resampleRecData(float *dat, float dt_prop, float maxFreq, bool keep_t0 = false, bool dipole = false,
                       float dz = 0.f, float vsurf = 0.f) {
    float w_pow = 0, phase_deg = 0, dt_spread = 1;
    if(dipole) w_pow = -1, phase_deg = 90, dt_spread = 0.5f * dz / vsurf;
    const int ntfft = klrtm_fft_size((int)nearbyintf(nt * 1.5f));
    klrtm_phasecorr *pc = nullptr;
    konglong::check(klrtm_phasecorr_create(&pc, nt, dt, dt_prop, -t0, maxFreq, ntfft, w_pow, phase_deg, 1, ntOut, dtOut,
                                           keep_t0 ? -t0 : 0.f));
    try {
      konglong::DeviceBuffer<float> in((size_t)nr * nt), out((size_t)nr * ntOut);
      in.upload(recDataRaw.data());
      konglong::check(klrtm_phasecorr_backward(pc, in.get(), nr, out.get(), 1 / dt_spread));
      out.download(dat);
    } catch(...) {
      klrtm_phasecorr_destroy(pc);
      throw;
    }
    klrtm_phasecorr_destroy(pc);
  }
  void saveToRecData(float dt_prop, float maxFreq, bool keep_t0 = false) {
    oData.resize((size_t)nr * ntOut);
    resampleRecData(oData.data(), dt_prop, maxFreq, keep_t0);
  }
  // the resampled record (saveToRecData first) with the receivers' coordinates [nr][3] (may be empty)
  void saveRecFile(const char *fileName, const std::vector<float> &xyz = {}) const {
    konglong::check(klrtm_record_write(fileName, nr, ntOut, dtOut, 0.f, xyz.empty() ? nullptr : xyz.data(),
                                       oData.data()));
  }

  int nt, ntOut;
  int nr;
  float dt, dtOut;
  float t0;
  float tmax;
  std::vector<float> recDataRaw;  // [nr][nt], host
  std::vector<float> oData;       // [nr][ntOut], host
};

#endif  // KONGLONG_EXTRACTRECEIVERDATA_H_
