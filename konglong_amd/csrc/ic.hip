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

// PREP: instead of correlating, write the tapered S and R to a.out_s / a.out_r (the operands of the
// FFT imaging conditions) and accumulate only the illumination
template<int FMT, bool RQ, bool PREP>
__global__ void ic_fused_kernel(const IcArgs a, const void *__restrict__ snap, const float *__restrict__ scales,
                                const float *__restrict__ w, const unsigned *__restrict__ rmax, float *__restrict__ img,
                                float *__restrict__ ill, double *__restrict__ img64, double *__restrict__ ill64) {
  // one y plane per blockIdx.y; a plane (nx*nz < 2^31 cells) is indexed with 32-bit arithmetic
  const unsigned nxz = (unsigned)a.nx * (unsigned)a.nz;
  const int iy = blockIdx.y;
  const float taper_t = ic_taper_t(a);
  if(a.step && a.snap_base) {  // the snapshot of this imaging step, picked on the device
    const int itsav = (*a.step / a.k) % a.nslots;
    snap = a.snap_base + a.snap_stride * (size_t)itsav;
    if(a.scales_base) scales = a.scales_base + a.nscales * (size_t)itsav;
  }
  for(unsigned j = blockIdx.x * blockDim.x + threadIdx.x; j < nxz; j += gridDim.x * blockDim.x) {
    const size_t i = (size_t)iy * nxz + j;
    const int ix = (int)(j / (unsigned)a.nz), iz = (int)(j - (unsigned)ix * (unsigned)a.nz);
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
    if(taper_t >= 0.f) S = S * taper_t;
    else if(iz < a.ntaper_z) S = S * a.taper_z[iz];
    float R = w[i];
    if(RQ) {
      float vm = __uint_as_float(rmax[iy]);
      R = vm > 1e-25f ? vm * c16_value(c16_shift(R * (1.0f / vm))) : 0.f;
    }
    R = edge_taper(a.T, R, ix, iy, iz, a.nx, a.ny, a.nz);
    if(PREP) {
      a.out_s[i] = S;
      a.out_r[i] = R;
    } else {
      const float d = S * R;
      if(img) img[i] = img[i] + d;
      if(img64) img64[i] = img64[i] + (double)S * (double)R;
    }
    if(ill) ill[i] = ill[i] + S * S;
    if(ill64) ill64[i] = ill64[i] + (double)S * (double)S;
  }
}

int launch_ic(int fmt, bool rq, const IcArgs &a, const void *snap, const float *scales, const float *w, const unsigned *rmax,
              float *img, float *ill, double *img64, double *ill64) {
  KL_REQUIRE((size_t)a.nx * a.nz < (1ull << 31) && a.ny <= 65535, KLRTM_EINVAL, "grid too large for the imaging kernel");
  const size_t nxz = (size_t)a.nx * a.nz;
  const dim3 grid((unsigned)std::min<size_t>((nxz + 255) / 256, 1024), (unsigned)a.ny);
  LaunchTimer tm("ic_correlate");
  const bool prep = a.out_s != nullptr;
  KL_REQUIRE(!prep || a.out_r, KLRTM_EINVAL, "imaging operands need both outputs");
#define KL_IC(F, Q)                                                                                                       \
  if(prep)                                                                                                                \
    hipLaunchKernelGGL((ic_fused_kernel<F, Q, true>), grid, dim3(256), 0, stream(), a, snap, scales, w, rmax, img, ill,   \
                       img64, ill64);                                                                                     \
  else                                                                                                                    \
    hipLaunchKernelGGL((ic_fused_kernel<F, Q, false>), grid, dim3(256), 0, stream(), a, snap, scales, w, rmax, img, ill,  \
                       img64, ill64)
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

// full[(y0+iy)*nxF + x0+ix][z0+iz] += img[iy][ix][iz] in FP64 (ImagingCondition::saveImage stacking,
// bgprtm/ImagingCondition.cpp:932-933, with FP64 accumulation across shots)
template<class T>
__global__ void stack_kernel(const T *__restrict__ img, int nx, int ny, int nz, double *__restrict__ full, int nxF,
                             int nzF, int x0, int y0, int z0) {
  const size_t n = (size_t)nx * ny * nz, nxz = (size_t)nx * nz;
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int iy = (int)(i / nxz);
    const size_t j = i - (size_t)iy * nxz;
    const int ix = (int)(j / nz), iz = (int)(j - (size_t)ix * nz);
    const size_t o = ((size_t)(y0 + iy) * nxF + (x0 + ix)) * nzF + (z0 + iz);
    full[o] = full[o] + (double)img[i];
  }
}

// the inline/crossline window copy of ImagingCondition::saveImage (bgprtm/ImagingCondition.cpp:874-945):
// rows il = iy0 .. iy1 (step incy), columns xl = ix0 .. ix1 (step incx) of the shot image go to the
// destination rows and columns, added (stack) or copied
__global__ void save_window_kernel(const float *__restrict__ in, float *__restrict__ out, int nz, int nyw, int nxw,
                                   size_t in_row, size_t in_col0, size_t out_row, size_t out_col0, int stack) {
  const size_t total = (size_t)nyw * nxw * nz;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int iz = (int)(e % nz);
    const size_t t = e / nz;
    const int jx = (int)(t % nxw), jy = (int)(t / nxw);
    const size_t i = ((size_t)jy * in_row + in_col0 + jx) * nz + iz, o = ((size_t)jy * out_row + out_col0 + jx) * nz + iz;
    out[o] = stack ? out[o] + in[i] : in[i];
  }
}

}  // namespace klrtm

using namespace klrtm;

extern "C" int klrtm_save_image(const float *img, int nz, int nx, int ny, int ix0s, int iy0s, int incx, int incy,
                                float *dest, int ix0d, int iy0d, int nxd, int nyd, int stack) {
  KL_REQUIRE(img && dest && nz > 0 && nx > 0 && ny > 0 && incx > 0 && incy > 0, KLRTM_EINVAL, "bad image");
  const int ix1s = ix0s + (nx - 1) * incx, iy1s = iy0s + (ny - 1) * incy;
  const int ix1d = ix0d + (nxd - 1) * incx, iy1d = iy0d + (nyd - 1) * incy;
  const int ix0 = std::max(ix0s, ix0d), iy0 = std::max(iy0s, iy0d), ix1 = std::min(ix1s, ix1d), iy1 = std::min(iy1s, iy1d);
  const int nyp = std::max(0, (iy1 - iy0) / incy + 1), nxp = std::max(0, (ix1 - ix0) / incx + 1);
  KL_REQUIRE((iy0s - iy0d) % incy == 0 && (ix0s - ix0d) % incx == 0, KLRTM_EINVAL,
             "input lines (%d, %d) and output lines (%d, %d) do not align on the increments (%d, %d)", iy0s, ix0s, iy0d,
             ix0d, incy, incx);
  KL_REQUIRE(!(nx == 1 && nxp == 0), KLRTM_EINVAL, "the crossline range does not meet the output grid");
  if(iy1 < iy0 || ix1 < ix0) return KLRTM_OK;  // no overlap
  const int iy0off = stack ? iy0d : iy0;
  const int ixoff = stack ? (ix0 - ix0d) / incx : 0;
  const int nxoff = stack ? nxd : nxp;
  const size_t total = (size_t)nyp * nxp * nz;
  const unsigned nb = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 256 * 64));
  // first row and column of the window in the input and in the output (rows of nx / nxoff columns)
  const float *in = img + (size_t)((iy0 - iy0s) / incy) * nx * nz;
  float *out = dest + (size_t)((iy0 - iy0off) / incy) * nxoff * nz;
  hipLaunchKernelGGL(save_window_kernel, dim3(nb), dim3(256), 0, stream(), in, out, nz, nyp, nxp, (size_t)nx,
                     (size_t)((ix0 - ix0s) / incx), (size_t)nxoff, (size_t)ixoff, stack);
  return check_launch("save_window_kernel");
}

extern "C" int klrtm_stack_image(const void *img, int is_double, int nx, int ny, int nz, double *full, int nxF, int nyF,
                                 int nzF, int x0, int y0, int z0) {
  KL_REQUIRE(img && full, KLRTM_EINVAL, "null image");
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0 && x0 >= 0 && y0 >= 0 && z0 >= 0 && x0 + nx <= nxF && y0 + ny <= nyF &&
                 z0 + nz <= nzF,
             KLRTM_EINVAL, "shot image %dx%dx%d at (%d,%d,%d) does not fit the %dx%dx%d output grid", nx, ny, nz, x0, y0,
             z0, nxF, nyF, nzF);
  const size_t n = (size_t)nx * ny * nz;
  const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
  LaunchTimer tm("stack");
  if(is_double)
    hipLaunchKernelGGL(stack_kernel<double>, dim3(nb), dim3(256), 0, stream(), (const double *)img, nx, ny, nz, full, nxF,
                       nzF, x0, y0, z0);
  else
    hipLaunchKernelGGL(stack_kernel<float>, dim3(nb), dim3(256), 0, stream(), (const float *)img, nx, ny, nz, full, nxF,
                       nzF, x0, y0, z0);
  return check_launch("stack_kernel");
}

// ImagingCondition::applyTape (bgprtm/ImagingCondition.cpp:200-226) in place
__global__ void apply_tape_kernel(float *__restrict__ wf, size_t nxy, int nz, float taper_t, const float *__restrict__ tz,
                                  int ntz) {
  const size_t tot = nxy * (size_t)(taper_t >= 0.f ? nz : ntz);
  const int lz = taper_t >= 0.f ? nz : ntz;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const size_t col = e / lz;
    const int iz = (int)(e - col * lz);
    const size_t i = col * nz + iz;
    wf[i] = wf[i] * (taper_t >= 0.f ? taper_t : tz[iz]);
  }
}

extern "C" int klrtm_apply_tape(float *wf, int nx, int ny, int nz, float taper_t, const float *taper_z, int ntaper_z) {
  KL_REQUIRE(wf && nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad wavefield");
  KL_REQUIRE(taper_t >= 0.f || ntaper_z <= 0 || taper_z, KLRTM_EINVAL, "z taper requested without taper values");
  if(taper_t < 0.f && ntaper_z <= 0) return KLRTM_OK;
  const size_t tot = (size_t)nx * ny * (taper_t >= 0.f ? nz : std::min(ntaper_z, nz));
  hipLaunchKernelGGL(apply_tape_kernel, dim3((unsigned)std::min<size_t>((tot + 255) / 256, 256 * 64)), dim3(256), 0, stream(),
                     wf, (size_t)nx * ny, nz, taper_t, taper_z, std::min(ntaper_z, nz));
  return check_launch("apply_tape_kernel");
}

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
