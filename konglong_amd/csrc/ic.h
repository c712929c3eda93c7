// ic.h -- imaging-condition launcher.
#pragma once
#include "klrtm_internal.h"
#include "codec_common.h"

namespace klrtm {
struct IcArgs {
  int nx, ny, nz;
  EdgeTaper T;
  float taper_t;         // >= 0: time taper factor of this snapshot; < 0: use the z taper
  const float *taper_z;  // device, ntaper_z values
  int ntaper_z;
  float *out_s = nullptr, *out_r = nullptr;  // set: write the tapered S and R here instead of correlating
};
struct Correlator;
int correlator_run(Correlator &C, float *image, float *img_fwi, float *img_raw, double *image64, float *src, float *rec,
                   int option);
int correlator_create(Correlator **out, int nx, int ny, int nz, float dx, float dy, float dz, float kmin);
int correlator_costerm(Correlator &C, float *image);
int correlator_sinterm(Correlator &C, float *imgfull, const float *imgcos);
void correlator_destroy(Correlator *C);
int launch_ic(int fmt, bool rq, const IcArgs &a, const void *snap, const float *scales, const float *w, const unsigned *rmax,
              float *img, float *ill, double *img64, double *ill64);
}  // namespace klrtm
