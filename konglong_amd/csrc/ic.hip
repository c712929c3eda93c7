// ic.hip -- zero-lag cross-correlation imaging condition.
//
// ImagingCondition::stackImageInterp/applyTape/calIllum and
// ImagingCorrelator::crossone (bgprtm/ImagingCondition.cpp:104-226,
// bgprtm/ImagingCorrelator.cpp:180-194):
//   S = decode(source snapshot) * (time taper | water-depth z taper)
//   R = decode(receiver snapshot)
//   image += S*R,  illum += S*S
// The reference reads both wavefields back from its snapshot caches; here the
// receiver side is the live backward wavefield, so the correlation runs inside
// the backward time loop and no receiver snapshots are stored.  When the
// reference's COMPRESS_16 cache is being reproduced, R is passed through the
// same 16-bit quantisation in registers (given the slice maxima of w0).
// FP64 outputs accumulate the same products in double (multi-shot stack).
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include "codec_common.h"
#include "ic.h"

namespace klrtm {

template<int FMT, bool RQ>
__global__ void ic_fused_kernel(const IcArgs a, const void *__restrict__ snap, const float *__restrict__ scales,
                                const float *__restrict__ w, const unsigned *__restrict__ rmax, float *__restrict__ img,
                                float *__restrict__ ill, double *__restrict__ img64, double *__restrict__ ill64) {
  const size_t nxz = (size_t)a.nx * a.nz;
  const size_t n = nxz * a.ny;
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int iy = (int)(i / nxz);
    const size_t j = i - (size_t)iy * nxz;
    const int ix = (int)(j / a.nz), iz = (int)(j - (size_t)ix * a.nz);
    float S;
    if(FMT == KLRTM_SNAP_FP32) {
      S = ((const float *)snap)[i];
    } else if(FMT == KLRTM_SNAP_C16) {
      const uint16_t *pin = (const uint16_t *)snap + (size_t)iy * (nxz + 2);
      float sc = __uint_as_float((unsigned)pin[0] | ((unsigned)pin[1] << 16));
      S = sc > 1e-25f ? sc * c16_value(pin[2 + j]) : 0.f;
    } else {
      S = half_decode<FMT>(((const uint16_t *)snap)[i]) * scales[block_index(ix, iy, iz, a.nx, a.nz)];
    }
    S = edge_taper(a.T, S, ix, iy, iz, a.nx, a.ny, a.nz);
    if(a.taper_t >= 0.f) S = S * a.taper_t;
    else if(iz < a.ntaper_z) S = S * a.taper_z[iz];
    float R = w[i];
    if(RQ) {
      float vm = __uint_as_float(rmax[iy]);
      R = vm > 1e-25f ? vm * c16_value(c16_shift(R * (1.0f / vm))) : 0.f;
    }
    R = edge_taper(a.T, R, ix, iy, iz, a.nx, a.ny, a.nz);
    const float d = S * R;
    if(img) img[i] = img[i] + d;
    if(ill) ill[i] = ill[i] + S * S;
    if(img64) img64[i] = img64[i] + (double)S * (double)R;
    if(ill64) ill64[i] = ill64[i] + (double)S * (double)S;
  }
}

int launch_ic(int fmt, bool rq, const IcArgs &a, const void *snap, const float *scales, const float *w, const unsigned *rmax,
              float *img, float *ill, double *img64, double *ill64) {
  size_t n = (size_t)a.nx * a.ny * a.nz;
  unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
  LaunchTimer tm("ic_correlate");
#define KL_IC(F, Q) \
  hipLaunchKernelGGL((ic_fused_kernel<F, Q>), dim3(nb), dim3(256), 0, stream(), a, snap, scales, w, rmax, img, ill, img64, ill64)
  if(fmt == KLRTM_SNAP_FP32) {
    if(rq) KL_IC(KLRTM_SNAP_FP32, true);
    else KL_IC(KLRTM_SNAP_FP32, false);
  } else if(fmt == KLRTM_SNAP_C16) {
    if(rq) KL_IC(KLRTM_SNAP_C16, true);
    else KL_IC(KLRTM_SNAP_C16, false);
  } else if(fmt == KLRTM_SNAP_FP16) {
    if(rq) KL_IC(KLRTM_SNAP_FP16, true);
    else KL_IC(KLRTM_SNAP_FP16, false);
  } else {
    if(rq) KL_IC(KLRTM_SNAP_BF16, true);
    else KL_IC(KLRTM_SNAP_BF16, false);
  }
#undef KL_IC
  return check_launch("ic_fused_kernel");
}

// standalone form: S and R already decoded; no edge taper (the caller decoded)
__global__ void ic_plain_kernel(const float *__restrict__ src, const float *__restrict__ rec, int nx, int ny, int nz,
                                float taper_t, const float *__restrict__ taper_z, int ntaper_z, float *__restrict__ img,
                                float *__restrict__ ill, double *__restrict__ img64, double *__restrict__ ill64) {
  const size_t n = (size_t)nx * ny * nz;
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    int iz = (int)(i % nz);
    float S = src[i];
    if(taper_t >= 0.f) S = S * taper_t;
    else if(iz < ntaper_z) S = S * taper_z[iz];
    float R = rec[i];
    float d = S * R;
    if(img) img[i] = img[i] + d;
    if(ill) ill[i] = ill[i] + S * S;
    if(img64) img64[i] = img64[i] + (double)S * (double)R;
    if(ill64) ill64[i] = ill64[i] + (double)S * (double)S;
  }
}

}  // namespace klrtm

using namespace klrtm;

extern "C" int klrtm_ic_correlate(const float *src, const float *rec, int nx, int ny, int nz, float taper_t,
                                  const float *taper_z, int ntaper_z, float *image, float *illum, double *image64,
                                  double *illum64) {
  KL_REQUIRE(src && rec, KLRTM_EINVAL, "null wavefield");
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad dimensions");
  KL_REQUIRE(taper_t >= 0.f || ntaper_z == 0 || taper_z, KLRTM_EINVAL, "z taper requested without taper values");
  size_t n = (size_t)nx * ny * nz;
  unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
  LaunchTimer tm("ic_correlate");
  hipLaunchKernelGGL(ic_plain_kernel, dim3(nb), dim3(256), 0, stream(), src, rec, nx, ny, nz, taper_t, taper_z, ntaper_z,
                     image, illum, image64, illum64);
  return check_launch("ic_plain_kernel");
}
