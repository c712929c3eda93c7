// surface.hip -- free-surface mirror of the wavefield, Propagator::apply_symmetry
// (bgprtm/Propagator.cpp:1857-1918), applied after every injection when the
// source or receiver side has a FREESURFACE boundary (bgprtm/Propagator.cpp:
// 1119-1125, 1228-1230).
//
// Above the surface index iz0 the column mirrors the field below it with sign
// sym (-1 for a pressure-release surface, which also zeroes the surface
// sample).  With the taper (taper_apply_symmetry, default on) the top tprlen
// samples (the z boundary) take a tapered, [1 4 6 4 1]/16-smoothed mirror
// image.  One lane per (x, y) column walks the column in the reference's
// order (the taper loop may read samples the first loop wrote when iz0 is
// close to tprlen), with the reference's mixed precision: the smoothing sum
// and its (1 - f) factor in FP32, the division by 16 and the rest in FP64,
// rounded to FP32 on the store.  FP contraction is off, so the result equals
// the sequential loop bit for bit.
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include "surface.h"
#include <cmath>
#include <vector>

namespace klrtm {

constexpr int SYM_MAX_TAPER = 256;
struct SymArgs {
  int nx, ny, nz, iz0, sym, tprlen, taper;
  float f[SYM_MAX_TAPER];  // taper weights 0.5*(1 - cos(pi*(iz + 0.5)/tprlen))
};

__global__ void symmetry_kernel(float *__restrict__ wf0, const SymArgs a) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if(c >= (long)a.nx * a.ny) return;
  float *wf = wf0 + (size_t)c * a.nz;
  const int iz2 = a.iz0 * 2;
  if(a.taper) {
    for(int iz = a.tprlen; iz < a.iz0; iz++) wf[iz] = a.sym * wf[iz2 - iz];
    if(a.sym == -1) wf[a.iz0] = 0;
    for(int iz = 0; iz < a.tprlen; iz++) {
      const float f = a.f[iz];
      const int m = iz2 - iz;
      const float sm = wf[m - 2] + wf[m + 2] + 4 * (wf[m - 1] + wf[m + 1]) + 6 * wf[m];
      const double t = (double)(wf[m] * f) + (double)((1 - f) * sm) / 16.;
      wf[iz] = (float)((a.sym * t) * (double)f);
    }
  } else {
    for(int iz = 0; iz < a.iz0; iz++) wf[iz] = a.sym * wf[iz2 - iz];
    if(a.sym == -1) wf[a.iz0] = 0;
    wf[0] = (float)(0.5 * (double)(wf[0] + wf[1]));
  }
}

int apply_symmetry_check(int nx, int ny, int nz, int iz0, int tprlen) {
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 1, KLRTM_EINVAL, "bad wavefield");
  KL_REQUIRE(iz0 >= 0 && 2 * iz0 < nz, KLRTM_EINVAL, "surface index %d needs 2*iz0 < nz (%d)", iz0, nz);
  KL_REQUIRE(tprlen == 0 || (tprlen > 0 && tprlen <= SYM_MAX_TAPER && tprlen <= iz0 && 2 * iz0 + 2 < nz), KLRTM_EINVAL,
             "taper length %d must lie in [0, min(%d, iz0 = %d)] with 2*iz0 + 2 < nz", tprlen, SYM_MAX_TAPER, iz0);
  return KLRTM_OK;
}

int apply_symmetry(float *wf, int nx, int ny, int nz, int iz0, int sym, int tprlen, int taper) {
  KL_REQUIRE(wf, KLRTM_EINVAL, "null wavefield");
  KL_REQUIRE(sym == 1 || sym == -1, KLRTM_EINVAL, "sym must be +1 or -1");
  int rc = apply_symmetry_check(nx, ny, nz, iz0, taper ? tprlen : 0);
  if(rc) return rc;
  SymArgs a{};
  a.nx = nx, a.ny = ny, a.nz = nz, a.iz0 = iz0, a.sym = sym, a.tprlen = tprlen, a.taper = taper;
  // the reference's promotions: (float)M_PI * (iz + 0.5) / tprlen in double, cosf of its float value
  for(int iz = 0; iz < tprlen && taper; iz++)
    a.f[iz] = 0.5f * (1 - cosf((float)((double)(float)M_PI * (iz + 0.5) / tprlen)));
  const long cols = (long)nx * ny;
  LaunchTimer tm("symmetry");
  hipLaunchKernelGGL(symmetry_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, stream(), wf, a);
  return check_launch("symmetry_kernel");
}

}  // namespace klrtm

extern "C" int klrtm_apply_symmetry(float *wf, int nx, int ny, int nz, int iz0, int sym, int tprlen, int taper) {
  return klrtm::apply_symmetry(wf, nx, ny, nz, iz0, sym, tprlen, taper);
}
