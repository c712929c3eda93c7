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
};
int launch_ic(int fmt, bool rq, const IcArgs &a, const void *snap, const float *scales, const float *w, const unsigned *rmax,
              float *img, float *ill, double *img64, double *ill64);
}  // namespace klrtm
