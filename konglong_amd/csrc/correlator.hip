// correlator.hip -- ImagingCorrelator's FFT imaging conditions on the device
// (bgprtm/ImagingCorrelator.cpp:17-431) at nDbl = 1 (no doubled volumes, so the
// libFFTV half-sample interpolation `interp` is the identity).
//
// Wavefields are padded per z trace to nz+2 floats and scaled by
// scalef = sqrt(1/(nx ny nz)) (buffPad2), transformed R2C in place with hipFFT,
// multiplied in the wavenumber domain and transformed back C2R; the unpadding
// scales by scalef again, so a round trip is the identity.
//   derive: H_a = i k_a / |k| for a = z, x, y (0 at k = 0), the wavenumbers of
//           assignW: -i dk for i <= n/2, (n - i) dk above;
//   inversion1: image += S R - Hz S Hz R - Hx S Hx R - Hy S Hy R, imgFWI += the
//           same products all with +1 (crossone with isign +1, -1, -1, -1);
//   inversion2: imgFWI += S R; S, R <- K S, K R with K = 1/max(|k|, kmin)
//           (kfilter, in place); image += S R;
//   allin3: imgRAW += |S R|, then inversion1;
//   costerm: image <- |k|^2 image (after the time loop of inversion2);
//   sinterm: imgfull <- 2 imgfull - 0.5 imgcos.
// The reference runs derive's three filters from one spectrum buffer that its
// 3D, ny2 == ny set-up aliases with the x and y outputs (srcBufx = srcBufy =
// srcBuff, bgprtm/ImagingCorrelator.cpp:42-51), so its lapy reads the x
// derivative in space as a spectrum; here every filter reads the intact
// spectrum, as the reference does when ny2 > ny.  The reference's in-place
// padded layout (stride nz+2) is FFTW's only for even nz, so nz must be even.
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include "ic.h"
#include <hipfft/hipfft.h>
#include <cmath>

namespace klrtm {

struct Correlator {
  int nx = 0, ny = 0, nz = 0;
  float dx = 0, dy = 0, dz = 0, kmin = 0;
  size_t nxy = 0, nxyz = 0, nzp2 = 0, padded = 0;  // padded = nxy * (nz + 2) floats per volume
  float scalef = 1.f;
  float *d_k = nullptr;      // kx[nx], ky[ny], kz[nz]
  float *spec = nullptr;     // two padded volumes: S, R (spectra after the forward transform)
  float *work = nullptr;     // two padded volumes: filtered spectra, then derivatives in space
  hipfftHandle r2c2 = 0, c2r2 = 0, r2c1 = 0, c2r1 = 0;  // batches of two and of one
  bool have[4] = {false, false, false, false};
  ~Correlator() {
    hipfftHandle *h[4] = {&r2c2, &c2r2, &r2c1, &c2r1};
    for(int i = 0; i < 4; i++)
      if(have[i]) (void)hipfftDestroy(*h[i]);
    for(void *p : {(void *)d_k, (void *)spec, (void *)work})
      if(p) (void)hipFree(p);
  }
};

#define KL_FFTC(call)                                                                      \
  do {                                                                                     \
    hipfftResult r_ = (call);                                                              \
    if(r_ != HIPFFT_SUCCESS) {                                                             \
      ::klrtm::set_error("%s failed: hipfft status %d (%s:%d)", #call, (int)r_, __FILE__, __LINE__); \
      return KLRTM_EDEVICE;                                                                \
    }                                                                                      \
  } while(0)

// buffPad2(+1): out[b][trace][iz] = in_b[trace][iz] * scalef, zero in the two pad samples
__global__ void corr_pad_kernel(const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ out,
                                size_t nxy, int nz, size_t padded, float scalef) {
  const size_t nzp2 = (size_t)nz + 2, tot = nxy * nzp2;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < 2 * tot; e += (size_t)gridDim.x * blockDim.x) {
    const int which = e >= tot;
    const size_t k = e - (which ? tot : 0);
    const size_t t = k / nzp2;
    const int iz = (int)(k - t * nzp2);
    const float *in = which ? b : a;
    float v = 0.f;
    if(in && iz < nz) v = in[t * nz + iz] * scalef;
    out[(which ? padded : 0) + k] = v;
  }
}

// buffPad2(-1): out[trace][iz] = in[trace][iz] * scalef
__global__ void corr_unpad_kernel(const float *__restrict__ in, float *__restrict__ out, size_t nxy, int nz, float scalef) {
  const size_t tot = nxy * nz;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const size_t t = e / nz;
    const int iz = (int)(e - t * nz);
    out[e] = in[t * ((size_t)nz + 2) + iz] * scalef;
  }
}

// |k| of bin (iy, ix, iz) in the reference's summation order
__device__ __forceinline__ float corr_kabs2(const float *k, int nx, int ny, int ix, int iy, int iz) {
  const float kx = k[ix], ky = k[nx + iy], kz = k[nx + ny + iz];
  return kx * kx + kz * kz + ky * ky;
}

// lapz / lapx / lapy (bgprtm/ImagingCorrelator.cpp:195-269) on both spectra: out = i (k_a/|k|) in
template<int AXIS>
__global__ void corr_unit_deriv_kernel(const float *__restrict__ in, float *__restrict__ out, const float *__restrict__ k,
                                       int nx, int ny, int nz, size_t padded) {
  const int nh = (nz + 2) / 2;
  const size_t tot = (size_t)nx * ny * nh;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < 2 * tot; e += (size_t)gridDim.x * blockDim.x) {
    const int which = e >= tot;
    const size_t c = e - (which ? tot : 0);
    const int iz = (int)(c % nh);
    const size_t t = c / nh;
    const int ix = (int)(t % nx), iy = (int)(t / nx);
    const float kk = sqrtf(corr_kabs2(k, nx, ny, ix, iy, iz));
    const size_t idd = (which ? padded : 0) + t * (size_t)(2 * nh) + 2 * iz;
    float re = 0.f, im = 0.f;
    if(kk > 1.0E-25f) {
      const float ka = AXIS == 0 ? k[nx + ny + iz] : AXIS == 1 ? k[ix] : k[nx + iy];
      const float r = ka / kk;
      im = r * in[idd];
      re = -r * in[idd + 1];
    }
    out[idd] = re;
    out[idd + 1] = im;
  }
}

// kfilter (bgprtm/ImagingCorrelator.cpp:285-306): spectrum *= 1/max(|k|, kmin), nv volumes
__global__ void corr_kfilter_kernel(float *__restrict__ s, const float *__restrict__ k, int nx, int ny, int nz,
                                    size_t padded, int nv, float kmin) {
  const int nh = (nz + 2) / 2;
  const size_t tot = (size_t)nx * ny * nh;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nv * tot; e += (size_t)gridDim.x * blockDim.x) {
    const int which = (int)(e / tot);
    const size_t c = e - which * tot;
    const int iz = (int)(c % nh);
    const size_t t = c / nh;
    const int ix = (int)(t % nx), iy = (int)(t / nx);
    float kk = sqrtf(corr_kabs2(k, nx, ny, ix, iy, iz));
    if(kk < kmin) kk = kmin;
    const float rkk = 1.0f / kk;
    const size_t idd = which * padded + t * (size_t)(2 * nh) + 2 * iz;
    s[idd] = s[idd] * rkk;
    s[idd + 1] = s[idd + 1] * rkk;
  }
}

// costerm (bgprtm/ImagingCorrelator.cpp:308-334): spectrum *= |k|^2
__global__ void corr_k2_kernel(float *__restrict__ s, const float *__restrict__ k, int nx, int ny, int nz) {
  const int nh = (nz + 2) / 2;
  const size_t tot = (size_t)nx * ny * nh;
  for(size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < tot; c += (size_t)gridDim.x * blockDim.x) {
    const int iz = (int)(c % nh);
    const size_t t = c / nh;
    const int ix = (int)(t % nx), iy = (int)(t / nx);
    const float kk = corr_kabs2(k, nx, ny, ix, iy, iz);
    const size_t idd = t * (size_t)(2 * nh) + 2 * iz;
    s[idd] = s[idd] * kk;
    s[idd + 1] = s[idd + 1] * kk;
  }
}

// crossone (bgprtm/ImagingCorrelator.cpp:180-194) of one term.  The operands are either the plain
// wavefields (PAD = false) or a padded pair from the C2R output, unpadded and scaled by scalef first
// (buffPad2(-1) into srcBuf[0]/recBuf[0]).  image += isign*dot, img2 += dot, raw += |dot|; FP64 image
// accumulates isign*dot in double.
template<bool PAD>
__global__ void corr_cross_kernel(const float *__restrict__ s, const float *__restrict__ r, size_t roff, size_t nxy, int nz,
                                  float scalef, float isign, float *__restrict__ image, float *__restrict__ img2,
                                  float *__restrict__ raw, double *__restrict__ image64) {
  const size_t tot = nxy * nz;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    float sv, rv;
    if(PAD) {
      const size_t t = e / nz, p = t * ((size_t)nz + 2) + (e - t * nz);
      sv = s[p] * scalef;
      rv = s[roff + p] * scalef;
    } else {
      sv = s[e];
      rv = r[e];
    }
    const float dot = sv * rv;
    if(raw) raw[e] = raw[e] + fabsf(dot);
    if(image) image[e] = image[e] + isign * dot;
    if(img2) img2[e] = img2[e] + dot;
    if(image64) image64[e] = image64[e] + (double)isign * ((double)sv * (double)rv);
  }
}

// imgmath (bgprtm/ImagingCorrelator.cpp:344-353)
__global__ void corr_imgmath_kernel(float *__restrict__ full, const float *__restrict__ imp, size_t n) {
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    full[e] = 2.0f * full[e] - 0.5f * imp[e];
}

static unsigned nblocks(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 256 * 64)); }

// assignW (bgprtm/ImagingCorrelator.cpp:141-150), dk = 2 pi / (d n) as the reference forms it
static void assign_w(int n, float d, float *w) {
  const float dw = (float)(2.0f * M_PI / (d * n));
  for(int i = 0; i < n / 2 + 1; i++) w[i] = -i * dw;
  for(int i = n / 2 + 1; i < n; i++) w[i] = (n - i) * dw;
}

static int plan(Correlator &C, hipfftHandle &h, bool &have, hipfftType type, int batch) {
  if(have) return KLRTM_OK;
  const int nh = (C.nz + 2) / 2;
  int n[3] = {C.ny, C.nx, C.nz}, re[3] = {C.ny, C.nx, C.nz + 2}, ce[3] = {C.ny, C.nx, nh};
  const int rank = C.ny > 1 ? 3 : 2;
  int *pn = rank == 3 ? n : n + 1, *pre = rank == 3 ? re : re + 1, *pce = rank == 3 ? ce : ce + 1;
  const int dist_r = (int)C.padded, dist_c = (int)(C.padded / 2);
  if(type == HIPFFT_R2C)
    KL_FFTC(hipfftPlanMany(&h, rank, pn, pre, 1, dist_r, pce, 1, dist_c, type, batch));
  else
    KL_FFTC(hipfftPlanMany(&h, rank, pn, pce, 1, dist_c, pre, 1, dist_r, type, batch));
  have = true;
  return KLRTM_OK;
}

static int exec(Correlator &C, hipfftHandle h, bool fwd, float *data) {
  KL_FFTC(hipfftSetStream(h, stream()));
  if(fwd) KL_FFTC(hipfftExecR2C(h, data, (hipfftComplex *)data));
  else KL_FFTC(hipfftExecC2R(h, (hipfftComplex *)data, data));
  return KLRTM_OK;
}

int correlator_create(Correlator **out, int nx, int ny, int nz, float dx, float dy, float dz, float kmin) {
  KL_REQUIRE(out, KLRTM_EINVAL, "null output");
  KL_REQUIRE(nx > 1 && ny > 0 && nz > 1, KLRTM_EINVAL, "bad dimensions %d x %d x %d", nx, ny, nz);
  KL_REQUIRE(nz % 2 == 0, KLRTM_EINVAL, "the padded in-place layout (nz+2 per trace) needs an even nz, got %d", nz);
  KL_REQUIRE(dx > 0 && dy > 0 && dz > 0 && kmin >= 0, KLRTM_EINVAL, "bad spacing or kmin");
  auto *C = new Correlator();
  C->nx = nx, C->ny = ny, C->nz = nz, C->dx = dx, C->dy = dy, C->dz = dz, C->kmin = kmin;
  C->nxy = (size_t)nx * ny, C->nxyz = C->nxy * nz, C->nzp2 = (size_t)nz + 2, C->padded = C->nxy * C->nzp2;
  if(C->padded >= (1ull << 31)) {
    delete C;
    set_error("volume of %zu padded samples exceeds the 32-bit FFT distance", (size_t)(nx * (size_t)ny * (nz + 2)));
    return KLRTM_EINVAL;
  }
  const float scaleX = 1.0f / float(nx), scaleY = 1.0f / float(ny), scaleZ = 1.0f / float(nz);
  C->scalef = sqrtf(scaleX * scaleY * scaleZ);
  std::vector<float> k(nx + ny + nz);
  assign_w(nx, dx, k.data());
  assign_w(ny, dy, k.data() + nx);
  assign_w(nz, dz, k.data() + nx + ny);
  if(hipMalloc(&C->d_k, sizeof(float) * k.size()) != hipSuccess || hipMalloc(&C->spec, sizeof(float) * 2 * C->padded) != hipSuccess ||
     hipMalloc(&C->work, sizeof(float) * 2 * C->padded) != hipSuccess) {
    delete C;
    set_error("correlator work space (4 padded volumes of %zu floats) does not fit", (size_t)(nx * (size_t)ny * (nz + 2)));
    return KLRTM_ENOMEM;
  }
  if(hipMemcpy(C->d_k, k.data(), sizeof(float) * k.size(), hipMemcpyHostToDevice) != hipSuccess) {
    delete C;
    set_error("correlator wavenumber upload failed");
    return KLRTM_EDEVICE;
  }
  *out = C;
  return KLRTM_OK;
}

// the padded, scaled spectra of a and b (either may be null: a zero volume) in C.spec
static int forward_pair(Correlator &C, const float *a, const float *b) {
  hipLaunchKernelGGL(corr_pad_kernel, dim3(nblocks(2 * C.padded)), dim3(256), 0, stream(), a, b, C.spec, C.nxy, C.nz,
                     C.padded, C.scalef);
  int rc = check_launch("corr_pad_kernel");
  if(rc || (rc = plan(C, C.r2c2, C.have[0], HIPFFT_R2C, 2))) return rc;
  return exec(C, C.r2c2, true, C.spec);
}

// derive + one signed crossone per axis (inversion1's derivative terms, in the order z, x, y)
static int derivative_terms(Correlator &C, float *image, float *img2, double *image64) {
  int rc = plan(C, C.c2r2, C.have[1], HIPFFT_C2R, 2);
  if(rc) return rc;
  const int nh = (C.nz + 2) / 2;
  const size_t ncplx = C.nxy * nh;
  for(int axis = 0; axis < (C.ny > 1 ? 3 : 2); axis++) {
    if(axis == 0)
      hipLaunchKernelGGL(corr_unit_deriv_kernel<0>, dim3(nblocks(2 * ncplx)), dim3(256), 0, stream(), C.spec, C.work,
                         C.d_k, C.nx, C.ny, C.nz, C.padded);
    else if(axis == 1)
      hipLaunchKernelGGL(corr_unit_deriv_kernel<1>, dim3(nblocks(2 * ncplx)), dim3(256), 0, stream(), C.spec, C.work,
                         C.d_k, C.nx, C.ny, C.nz, C.padded);
    else
      hipLaunchKernelGGL(corr_unit_deriv_kernel<2>, dim3(nblocks(2 * ncplx)), dim3(256), 0, stream(), C.spec, C.work,
                         C.d_k, C.nx, C.ny, C.nz, C.padded);
    if((rc = check_launch("corr_unit_deriv_kernel")) || (rc = exec(C, C.c2r2, false, C.work))) return rc;
    hipLaunchKernelGGL(corr_cross_kernel<true>, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), C.work, nullptr, C.padded,
                       C.nxy, C.nz, C.scalef, -1.0f, image, img2, nullptr, image64);
    if((rc = check_launch("corr_cross_kernel"))) return rc;
  }
  return KLRTM_OK;
}

static int cross_plain(Correlator &C, const float *s, const float *r, float isign, float *image, float *img2, float *raw,
                       double *image64) {
  hipLaunchKernelGGL(corr_cross_kernel<false>, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), s, r, (size_t)0, C.nxy,
                     C.nz, C.scalef, isign, image, img2, raw, image64);
  return check_launch("corr_cross_kernel");
}

// K applied in place to nv volumes (nv = 1: a, nv = 2: a and b)
static int kfilter(Correlator &C, float *a, float *b) {
  int rc = forward_pair(C, a, b);
  if(rc || (rc = plan(C, C.c2r2, C.have[1], HIPFFT_C2R, 2))) return rc;
  const size_t ncplx = C.nxy * ((C.nz + 2) / 2);
  hipLaunchKernelGGL(corr_kfilter_kernel, dim3(nblocks(2 * ncplx)), dim3(256), 0, stream(), C.spec, C.d_k, C.nx, C.ny,
                     C.nz, C.padded, 2, C.kmin);
  if((rc = check_launch("corr_kfilter_kernel")) || (rc = exec(C, C.c2r2, false, C.spec))) return rc;
  hipLaunchKernelGGL(corr_unpad_kernel, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), C.spec, a, C.nxy, C.nz, C.scalef);
  if((rc = check_launch("corr_unpad_kernel"))) return rc;
  if(b) {
    hipLaunchKernelGGL(corr_unpad_kernel, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), C.spec + C.padded, b, C.nxy, C.nz,
                       C.scalef);
    rc = check_launch("corr_unpad_kernel");
  }
  return rc;
}

int correlator_run(Correlator &C, float *image, float *img_fwi, float *img_raw, double *image64, float *src, float *rec,
                   int option) {
  KL_REQUIRE(src && rec, KLRTM_EINVAL, "null wavefield");
  LaunchTimer tm("ic_correlator");
  int rc = KLRTM_OK;
  switch(option) {
  case KLRTM_IC_IMAGE1: return cross_plain(C, src, rec, 1.0f, image, nullptr, nullptr, image64);
  case KLRTM_IC_ALLIN3:
    if(img_raw && (rc = cross_plain(C, src, rec, 1.0f, nullptr, nullptr, img_raw, nullptr))) return rc;
    if(!image && !img_fwi && !image64) return KLRTM_OK;
    // fall through: inversion1 (bgprtm/ImagingCorrelator.cpp:428-434)
  case KLRTM_IC_INVERSION:
    if((rc = forward_pair(C, src, rec))) return rc;
    if((rc = cross_plain(C, src, rec, 1.0f, image, img_fwi, nullptr, image64))) return rc;
    return derivative_terms(C, image, img_fwi, image64);
  case KLRTM_IC_INVERSION2:
    if(img_fwi && (rc = cross_plain(C, src, rec, 1.0f, img_fwi, nullptr, nullptr, nullptr))) return rc;
    if((rc = kfilter(C, src, rec))) return rc;
    return cross_plain(C, src, rec, 1.0f, image, nullptr, nullptr, image64);
  default: set_error("unknown imaging option %d", option); return KLRTM_EINVAL;
  }
}

int correlator_costerm(Correlator &C, float *image) {
  KL_REQUIRE(image, KLRTM_EINVAL, "null image");
  int rc;
  hipLaunchKernelGGL(corr_pad_kernel, dim3(nblocks(2 * C.padded)), dim3(256), 0, stream(), image, nullptr, C.work, C.nxy,
                     C.nz, C.padded, C.scalef);
  if((rc = check_launch("corr_pad_kernel")) || (rc = plan(C, C.r2c1, C.have[2], HIPFFT_R2C, 1)) ||
     (rc = plan(C, C.c2r1, C.have[3], HIPFFT_C2R, 1)) || (rc = exec(C, C.r2c1, true, C.work)))
    return rc;
  hipLaunchKernelGGL(corr_k2_kernel, dim3(nblocks(C.nxy * ((C.nz + 2) / 2))), dim3(256), 0, stream(), C.work, C.d_k, C.nx,
                     C.ny, C.nz);
  if((rc = check_launch("corr_k2_kernel")) || (rc = exec(C, C.c2r1, false, C.work))) return rc;
  hipLaunchKernelGGL(corr_unpad_kernel, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), C.work, image, C.nxy, C.nz, C.scalef);
  return check_launch("corr_unpad_kernel");
}

int correlator_sinterm(Correlator &C, float *imgfull, const float *imgcos) {
  if(!imgfull) return KLRTM_OK;
  KL_REQUIRE(imgcos, KLRTM_EINVAL, "null cos-term image");
  hipLaunchKernelGGL(corr_imgmath_kernel, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), imgfull, imgcos, C.nxyz);
  return check_launch("corr_imgmath_kernel");
}

// H_axis(in) -> out (axis 0 z, 1 x, 2 y), as derive leaves it in srcBuf[0] after buffPad2(-1)
int correlator_unit_derivative(Correlator &C, const float *in, int axis, float *out) {
  KL_REQUIRE(in && out && axis >= 0 && axis <= 2, KLRTM_EINVAL, "bad derivative request");
  KL_REQUIRE(axis < 2 || C.ny > 1, KLRTM_EINVAL, "no y derivative on a 2D grid");
  int rc = forward_pair(C, in, nullptr);
  if(rc || (rc = plan(C, C.c2r2, C.have[1], HIPFFT_C2R, 2))) return rc;
  const size_t ncplx = C.nxy * ((C.nz + 2) / 2);
  if(axis == 0)
    hipLaunchKernelGGL(corr_unit_deriv_kernel<0>, dim3(nblocks(2 * ncplx)), dim3(256), 0, stream(), C.spec, C.work, C.d_k,
                       C.nx, C.ny, C.nz, C.padded);
  else if(axis == 1)
    hipLaunchKernelGGL(corr_unit_deriv_kernel<1>, dim3(nblocks(2 * ncplx)), dim3(256), 0, stream(), C.spec, C.work, C.d_k,
                       C.nx, C.ny, C.nz, C.padded);
  else
    hipLaunchKernelGGL(corr_unit_deriv_kernel<2>, dim3(nblocks(2 * ncplx)), dim3(256), 0, stream(), C.spec, C.work, C.d_k,
                       C.nx, C.ny, C.nz, C.padded);
  if((rc = check_launch("corr_unit_deriv_kernel")) || (rc = exec(C, C.c2r2, false, C.work))) return rc;
  hipLaunchKernelGGL(corr_unpad_kernel, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), C.work, out, C.nxy, C.nz, C.scalef);
  return check_launch("corr_unpad_kernel");
}

void correlator_destroy(Correlator *C) { delete C; }

int correlator_kfilter(Correlator &C, float *wave) {
  KL_REQUIRE(wave, KLRTM_EINVAL, "null wavefield");
  return kfilter(C, wave, nullptr);
}

}  // namespace klrtm

using namespace klrtm;

extern "C" {
int klrtm_correlator_create(klrtm_correlator **c, int nx, int ny, int nz, float dx, float dy, float dz, float kmin) {
  Correlator *C = nullptr;
  int rc = correlator_create(&C, nx, ny, nz, dx, dy, dz, kmin);
  if(rc == KLRTM_OK) *c = reinterpret_cast<klrtm_correlator *>(C);
  return rc;
}
int klrtm_correlator_destroy(klrtm_correlator *c) {
  correlator_destroy(reinterpret_cast<Correlator *>(c));
  return KLRTM_OK;
}
int klrtm_correlator_run(klrtm_correlator *c, float *image, float *img_fwi, float *img_raw, double *image64, float *src,
                         float *rec, int option) {
  KL_REQUIRE(c, KLRTM_EINVAL, "null correlator");
  return correlator_run(*reinterpret_cast<Correlator *>(c), image, img_fwi, img_raw, image64, src, rec, option);
}
int klrtm_correlator_costerm(klrtm_correlator *c, float *image) {
  KL_REQUIRE(c, KLRTM_EINVAL, "null correlator");
  return correlator_costerm(*reinterpret_cast<Correlator *>(c), image);
}
int klrtm_correlator_sinterm(klrtm_correlator *c, float *imgfull, const float *imgcos) {
  KL_REQUIRE(c, KLRTM_EINVAL, "null correlator");
  return correlator_sinterm(*reinterpret_cast<Correlator *>(c), imgfull, imgcos);
}
int klrtm_correlator_unit_derivative(klrtm_correlator *c, const float *in, int axis, float *out) {
  KL_REQUIRE(c, KLRTM_EINVAL, "null correlator");
  return correlator_unit_derivative(*reinterpret_cast<Correlator *>(c), in, axis, out);
}
int klrtm_correlator_kfilter(klrtm_correlator *c, float *wave) {
  KL_REQUIRE(c, KLRTM_EINVAL, "null correlator");
  return correlator_kfilter(*reinterpret_cast<Correlator *>(c), wave);
}
}
