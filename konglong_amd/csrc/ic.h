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
  // set (captured step sequences): the snapshot index itsav = (*step / k) % nslots is read on the device,
  // which picks the time taper taper_tab[itsav] (itsav < ntaper_t, else the z taper) and, for the
  // launchers that take them, the snapshot snap_base + itsav*snap_stride and its block scales
  const int *step = nullptr;
  int k = 1, nslots = 1, ntaper_t = 0;
  const float *taper_tab = nullptr;
  const char *snap_base = nullptr;
  size_t snap_stride = 0;
  const float *scales_base = nullptr;
  size_t nscales = 0;
};
// the time taper factor of an imaging launch (< 0: use the z taper)
__device__ __forceinline__ float ic_taper_t(const IcArgs &a) {
  if(!a.step) return a.taper_t;
  const int itsav = (*a.step / a.k) % a.nslots;
  return itsav < a.ntaper_t ? a.taper_tab[itsav] : -1.f;
}
struct Correlator;
int correlator_run(Correlator &C, float *image, float *img_fwi, float *img_raw, double *image64, float *src, float *rec,
                   int option);
int correlator_create(Correlator **out, int nx, int ny, int nz, float dx, float dy, float dz, float kmin);
int correlator_costerm(Correlator &C, float *image);
int correlator_sinterm(Correlator &C, float *imgfull, const float *imgcos);
void correlator_destroy(Correlator *C);
// doubled imaging grid (factors 1 or 2 per axis z, x, y; high-cut fractions of the Nyquist wavenumber)
int correlator_set_double(Correlator &C, int dz, int dx, int dy, float high_z, float high_xy);
int correlator_nloop(const Correlator &C);
int correlator_interleave(Correlator &C, const float *sub, float *out);
int correlator_add_illum(Correlator &C, const float *s, float *ill);
int correlator_run_copies(Correlator &C, float *image, float *img_fwi, float *src, float *rec, int option);
int correlator_interp_illum(Correlator &C, const float *in, float *out);
int launch_ic(int fmt, bool rq, const IcArgs &a, const void *snap, const float *scales, const float *w, const unsigned *rmax,
              float *img, float *ill, double *img64, double *ill64);
}  // namespace klrtm
