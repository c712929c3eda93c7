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
// Doubled imaging grids (ic_double_z/x/y = 2, bgprtm/ImagingCondition.cpp:61-67, 104-141, 169-178,
// 415-431; bgprtm/ImagingCorrelator.cpp:58-75, 353-524): S and R are high-cut along every doubled axis
// (HIGHTAPE: highTapeZ 0.9 along z, highTapeXY 0.7 along x and y), and before every correlation their
// half-sample shifted copies are formed along the doubled axes in the reference's order (z, x, y, zx,
// zy, xy, zxy from copy 0); copy i of S times copy i of R accumulates into sub-image i, and interleave
// places sub-image i at its parity on the doubled grid.  Both filters are libFFTV's, which is not
// vendored: here they are defined on zero-padded 1D FFTs along the axis (hipFFT, length the smallest
// 2^a 3^b 5^c >= n + 16): the shift multiplies bin k of length L by exp(i pi k / L) (band-limited
// interpolation at x + 1/2, the Nyquist bin dropping out), the high cut by 1 up to h times the Nyquist
// wavenumber and a cosine to 0 at it.
// The reference runs derive's three filters from one spectrum buffer that its
// 3D, ny2 == ny set-up aliases with the x and y outputs (srcBufx = srcBufy =
// srcBuff, bgprtm/ImagingCorrelator.cpp:42-51), so its lapy reads the x
// derivative in space as a spectrum; here every filter reads the intact
// spectrum, as the reference does when ny2 > ny.  The reference's in-place
// padded layout (stride nz+2) is FFTW's only for even nz, so nz must be even.
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include "ic.h"
#include "timeshift.h"
#include <hipfft/hipfft.h>
#include <algorithm>
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
  // doubled grid: factors per axis (z, x, y), copies, the copies' shift bits, copy buffers of S and R
  // (nloop volumes each), high-cut fractions, and the 1D line transforms per axis
  int dbl[3] = {1, 1, 1};
  int nloop = 1;
  int bits[8] = {0};  // copy i: bit 0 z shift, bit 1 x shift, bit 2 y shift
  int src_of[8] = {0};  // copy i is copy src_of[i] shifted along axis_of[i]
  int axis_of[8] = {0};
  float high_z = 0.9f, high_xy = 0.7f;
  float *sbuf = nullptr, *rbuf = nullptr, *lines = nullptr;
  int L[3] = {0, 0, 0};
  hipfftHandle lr2c[3] = {0, 0, 0}, lc2r[3] = {0, 0, 0};
  bool lhave[3] = {false, false, false};
  ~Correlator() {
    hipfftHandle *h[4] = {&r2c2, &c2r2, &r2c1, &c2r1};
    for(int i = 0; i < 4; i++)
      if(have[i]) (void)hipfftDestroy(*h[i]);
    for(int a = 0; a < 3; a++)
      if(lhave[a]) (void)hipfftDestroy(lr2c[a]), (void)hipfftDestroy(lc2r[a]);
    for(void *p : {(void *)d_k, (void *)spec, (void *)work, (void *)sbuf, (void *)rbuf, (void *)lines})
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

// Lines along axis a (0 z, 1 x, 2 y) of a volume [ny][nx][nz] copied into rows of Lp = L + 2 floats
// (zero beyond n_a: the transform's padding); line index = the other two coordinates in memory order.
__global__ void corr_lines_gather_kernel(const float *__restrict__ in, float *__restrict__ lines, int nx, int ny, int nz,
                                         int axis, int L) {
  const int na = axis == 0 ? nz : axis == 1 ? nx : ny;
  const size_t nl = (size_t)nx * ny * nz / na, Lp = (size_t)L + 2, tot = nl * Lp;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const size_t l = e / Lp;
    const int j = (int)(e - l * Lp);
    float v = 0.f;
    if(j < na) {
      size_t idx;
      if(axis == 0) idx = l * nz + j;
      else if(axis == 1) idx = ((l / nz) * nx + j) * nz + l % nz;  // l = iy*nz + iz
      else idx = ((size_t)j * nx) * nz + l;                        // l = ix*nz + iz
      v = in[idx];
    }
    lines[e] = v;
  }
}
__global__ void corr_lines_scatter_kernel(const float *__restrict__ lines, float *__restrict__ out, int nx, int ny, int nz,
                                          int axis, int L, float scale) {
  const int na = axis == 0 ? nz : axis == 1 ? nx : ny;
  const size_t nl = (size_t)nx * ny * nz / na, Lp = (size_t)L + 2, tot = nl * (size_t)na;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const size_t l = e / na;
    const int j = (int)(e - l * na);
    size_t idx;
    if(axis == 0) idx = l * nz + j;
    else if(axis == 1) idx = ((l / nz) * nx + j) * nz + l % nz;
    else idx = ((size_t)j * nx) * nz + l;
    out[idx] = lines[l * Lp + j] * scale;
  }
}
// per-bin factor of the line spectra: half-sample shift (mode 0) or high cut at fraction h (mode 1)
__global__ void corr_lines_filter_kernel(float *__restrict__ lines, size_t nl, int L, int mode, float h) {
  const int nb = L / 2 + 1;
  const size_t tot = nl * nb, Lp = (size_t)L + 2;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const size_t l = e / nb;
    const int k = (int)(e - l * nb);
    float *c = lines + l * Lp + 2 * k;
    if(mode == 0) {
      float sn, cs;
      sincospif((float)k / (float)L, &sn, &cs);
      const float re = c[0] * cs - c[1] * sn, im = c[0] * sn + c[1] * cs;
      c[0] = re, c[1] = im;
    } else {
      const float f = (float)k / (0.5f * L);
      const float w = f <= h ? 1.f : 0.5f * (1.f + cospif((f - h) / (1.f - h)));
      c[0] *= w, c[1] *= w;
    }
  }
}
// x and y lines through 32 x 32 LDS tiles: the volume side is read / written along the line index (z fastest,
// contiguous) and the padded rows along the sample index, so both sides are coalesced.  Tile (lt, jt) covers
// lines 32 lt.. and samples 32 jt..; a block is 32 x 8 lanes, each moving four elements.
__device__ __forceinline__ size_t xline_index(int axis, size_t l, int j, int nx, int nz, size_t nl) {
  return axis == 1 ? ((l / nz) * nx + j) * (size_t)nz + l % nz : (size_t)j * nl + l;
}
__global__ void __launch_bounds__(256) corr_xlines_gather_kernel(const float *__restrict__ in, float *__restrict__ lines,
                                                                 int nx, int ny, int nz, int axis, int L) {
  __shared__ float t[32][33];
  const int na = axis == 1 ? nx : ny, Lp = L + 2;
  const size_t nl = (size_t)nx * ny * nz / na;
  const size_t l0 = (size_t)blockIdx.x * 32;
  const int j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for(int r = 0; r < 4; r++) {  // read: lanes along l, rows along j
    const int j = j0 + ty + 8 * r;
    const size_t l = l0 + tx;
    t[ty + 8 * r][tx] = (l < nl && j < na) ? in[xline_index(axis, l, j, nx, nz, nl)] : 0.f;
  }
  __syncthreads();
#pragma unroll
  for(int r = 0; r < 4; r++) {  // write: lanes along j, rows along l
    const size_t l = l0 + ty + 8 * r;
    const int j = j0 + tx;
    if(l < nl && j < Lp) lines[l * Lp + j] = t[tx][ty + 8 * r];
  }
}
__global__ void __launch_bounds__(256) corr_xlines_scatter_kernel(const float *__restrict__ lines, float *__restrict__ out,
                                                                  int nx, int ny, int nz, int axis, int L, float scale) {
  __shared__ float t[32][33];
  const int na = axis == 1 ? nx : ny, Lp = L + 2;
  const size_t nl = (size_t)nx * ny * nz / na;
  const size_t l0 = (size_t)blockIdx.x * 32;
  const int j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for(int r = 0; r < 4; r++) {  // read: lanes along j
    const size_t l = l0 + ty + 8 * r;
    const int j = j0 + tx;
    t[ty + 8 * r][tx] = (l < nl && j < na) ? lines[l * Lp + j] * scale : 0.f;
  }
  __syncthreads();
#pragma unroll
  for(int r = 0; r < 4; r++) {  // write: lanes along l
    const int j = j0 + ty + 8 * r;
    const size_t l = l0 + tx;
    if(l < nl && j < na) out[xline_index(axis, l, j, nx, nz, nl)] = t[tx][ty + 8 * r];
  }
}

// sub-images [nloop][ny][nx][nz] -> the doubled grid [ny*dy][nx*dx][nz*dz]: copy i at its shift parity
__global__ void corr_interleave_kernel(const float *__restrict__ sub, float *__restrict__ out, int nx, int ny, int nz,
                                       int dz, int dx, int dy, int nloop, const int4 bits03, const int4 bits47) {
  const size_t n = (size_t)nx * ny * nz, tot = n * nloop;
  const int b[8] = {bits03.x, bits03.y, bits03.z, bits03.w, bits47.x, bits47.y, bits47.z, bits47.w};
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n);
    const size_t c = e - (size_t)i * n;
    const int iz = (int)(c % nz);
    const size_t t = c / nz;
    const int ix = (int)(t % nx), iy = (int)(t / nx);
    const int oz = iz * dz + (b[i] & 1), ox = ix * dx + ((b[i] >> 1) & 1), oy = iy * dy + ((b[i] >> 2) & 1);
    out[((size_t)oy * nx * dx + ox) * ((size_t)nz * dz) + oz] = sub[e];
  }
}

// ill += s * s (ImagingCondition::calIllum on the high-cut source wavefield, bgprtm/ImagingCondition.cpp:143-149)
__global__ void corr_illum_kernel(const float *__restrict__ s, float *__restrict__ ill, size_t n) {
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    ill[i] += s[i] * s[i];
}
// the illumination onto the doubled grid by averages of the 2, 4 or 8 surrounding samples
// (ImagingCondition::interpIllum, bgprtm/ImagingCondition.cpp:759-810): one thread per coarse point writes
// its eight targets in the reference's order and sums in its order; on an undoubled axis the "second"
// target coincides with the first and the later write stands, as there
__global__ void corr_interp_illum_kernel(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nz,
                                         int dz, int dx, int dy) {
  const size_t n = (size_t)nx * ny * nz, nx2 = (size_t)nx * dx, nz2 = (size_t)nz * dz;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int iz = (int)(e % nz);
    const size_t t = e / nz;
    const int ix = (int)(t % nx), iy = (int)(t / nx);
    const int y1 = iy * dy, y2 = iy * dy + dy - 1, x1 = ix * dx, x2 = ix * dx + dx - 1, z1 = iz * dz, z2 = iz * dz + dz - 1;
    const int yp = min(iy + 1, ny - 1), xp = min(ix + 1, nx - 1), zp = min(iz + 1, nz - 1);
    auto at = [&](int y, int x, int z) { return in[((size_t)y * nx + x) * nz + z]; };
    auto o = [&](int y, int x, int z) -> float & { return out[((size_t)y * nx2 + x) * nz2 + z]; };
    const float c = at(iy, ix, iz);
    o(y1, x1, z1) = c;
    o(y2, x1, z1) = 0.5f * (c + at(yp, ix, iz));
    o(y1, x2, z1) = 0.5f * (c + at(iy, xp, iz));
    o(y1, x1, z2) = 0.5f * (c + at(iy, ix, zp));
    o(y2, x1, z2) = 0.25f * (c + at(iy, ix, zp) + at(yp, ix, iz) + at(yp, ix, zp));
    o(y1, x2, z2) = 0.25f * (c + at(iy, ix, zp) + at(iy, xp, iz) + at(iy, xp, zp));
    o(y2, x2, z1) = 0.25f * (c + at(yp, ix, iz) + at(iy, xp, iz) + at(yp, xp, iz));
    o(y2, x2, z2) = 0.125f * (c + at(yp, ix, iz) + at(iy, xp, iz) + at(iy, ix, zp) + at(yp, ix, zp) + at(yp, xp, iz) +
                              at(iy, xp, zp) + at(yp, xp, zp));
  }
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

// ---- doubled grids ---------------------------------------------------------------------------------

// in place, one padded row of L + 2 floats per line
static int line_plans(Correlator &C, int axis) {
  if(C.lhave[axis]) return KLRTM_OK;
  const int na = axis == 0 ? C.nz : axis == 1 ? C.nx : C.ny;
  const int Lx = C.L[axis], nb = Lx / 2 + 1, Lp = Lx + 2;
  const int nl = (int)(C.nxyz / na);
  int n[1] = {Lx}, re[1] = {Lp}, ce[1] = {nb};
  KL_FFTC(hipfftPlanMany(&C.lr2c[axis], 1, n, re, 1, Lp, ce, 1, nb, HIPFFT_R2C, nl));
  KL_FFTC(hipfftPlanMany(&C.lc2r[axis], 1, n, ce, 1, nb, re, 1, Lp, HIPFFT_C2R, nl));
  C.lhave[axis] = true;
  return KLRTM_OK;
}

// out = the axis filter of in (mode 0: half-sample shift, 1: high cut at h); in == out allowed
static int axis_filter(Correlator &C, const float *in, float *out, int axis, int mode, float h) {
  int rc = line_plans(C, axis);
  if(rc) return rc;
  const int na = axis == 0 ? C.nz : axis == 1 ? C.nx : C.ny, Lx = C.L[axis];
  const size_t nl = C.nxyz / na;
  const unsigned lt = (unsigned)((nl + 31) / 32);
  if(axis > 0)
    hipLaunchKernelGGL(corr_xlines_gather_kernel, dim3(lt, (unsigned)((Lx + 2 + 31) / 32)), dim3(256), 0, stream(), in,
                       C.lines, C.nx, C.ny, C.nz, axis, Lx);
  else
    hipLaunchKernelGGL(corr_lines_gather_kernel, dim3(nblocks(nl * (Lx + 2))), dim3(256), 0, stream(), in, C.lines, C.nx,
                       C.ny, C.nz, axis, Lx);
  KL_FFTC(hipfftSetStream(C.lr2c[axis], stream()));
  KL_FFTC(hipfftExecR2C(C.lr2c[axis], C.lines, (hipfftComplex *)C.lines));
  hipLaunchKernelGGL(corr_lines_filter_kernel, dim3(nblocks(nl * (Lx / 2 + 1))), dim3(256), 0, stream(), C.lines, nl, Lx,
                     mode, h);
  KL_FFTC(hipfftSetStream(C.lc2r[axis], stream()));
  KL_FFTC(hipfftExecC2R(C.lc2r[axis], (hipfftComplex *)C.lines, C.lines));
  if(axis > 0)
    hipLaunchKernelGGL(corr_xlines_scatter_kernel, dim3(lt, (unsigned)((na + 31) / 32)), dim3(256), 0, stream(), C.lines,
                       out, C.nx, C.ny, C.nz, axis, Lx, 1.0f / (float)Lx);
  else
    hipLaunchKernelGGL(corr_lines_scatter_kernel, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), C.lines, out, C.nx,
                       C.ny, C.nz, axis, Lx, 1.0f / (float)Lx);
  return check_launch("corr_lines_scatter_kernel");
}

// ImagingCondition::applyFFTV: high cut along the doubled axes (z, x, y), in place
static int hightape(Correlator &C, float *v) {
  int rc = KLRTM_OK;
  for(int a = 0; a < 3 && !rc; a++)
    if(C.dbl[a] > 1) rc = axis_filter(C, v, v, a, 1, a == 0 ? C.high_z : C.high_xy);
  return rc;
}

// copy 0 of buf (nloop volumes) holds the field; copies 1.. are its half-sample shifts
// (ImagingCorrelator::interp, bgprtm/ImagingCorrelator.cpp:515-524)
static int interp(Correlator &C, float *buf) {
  int rc = KLRTM_OK;
  for(int i = 1; i < C.nloop && !rc; i++)
    rc = axis_filter(C, buf + (size_t)C.src_of[i] * C.nxyz, buf + (size_t)i * C.nxyz, C.axis_of[i], 0, 0.f);
  return rc;
}

// one signed correlation term per sub-image of the nloop-volume buffers sb, rb whose copy 0 holds the fields:
// copies 1.. are formed in place, then sub-image i takes copy i of S times copy i of R (ImagingCorrelator's
// interp + crossone loop, bgprtm/ImagingCorrelator.cpp:374-414)
static int cross_copies(Correlator &C, float *sb, float *rb, float isign, float *image, float *img2) {
  int rc;
  if((rc = interp(C, sb)) || (rc = interp(C, rb))) return rc;
  for(int i = 0; i < C.nloop && !rc; i++) {
    const size_t o = (size_t)i * C.nxyz;
    rc = cross_plain(C, sb + o, rb + o, isign, image ? image + o : nullptr, img2 ? img2 + o : nullptr, nullptr, nullptr);
  }
  return rc;
}

// the same for (copy-0) fields s and r held elsewhere: copied into the correlator's copy buffers first
static int cross_term(Correlator &C, const float *s, const float *r, float isign, float *image, float *img2) {
  KL_HIP(hipMemcpyAsync(C.sbuf, s, C.nxyz * sizeof(float), hipMemcpyDeviceToDevice, stream()));
  KL_HIP(hipMemcpyAsync(C.rbuf, r, C.nxyz * sizeof(float), hipMemcpyDeviceToDevice, stream()));
  return cross_copies(C, C.sbuf, C.rbuf, isign, image, img2);
}

// inversion1's derivative terms on the doubled grid from the spectra in C.spec: each H_a S, H_a R into copy
// 0 of sb, rb (as buffPad2(-1) leaves them in srcBuf[0]), then cross_copies
static int derivative_terms_dbl(Correlator &C, float *sb, float *rb, float *image, float *img2) {
  int rc = plan(C, C.c2r2, C.have[1], HIPFFT_C2R, 2);
  if(rc) return rc;
  const size_t ncplx = C.nxy * ((C.nz + 2) / 2);
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
    hipLaunchKernelGGL(corr_unpad_kernel, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), C.work, sb, C.nxy, C.nz,
                       C.scalef);
    hipLaunchKernelGGL(corr_unpad_kernel, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), C.work + C.padded, rb, C.nxy,
                       C.nz, C.scalef);
    if((rc = check_launch("corr_unpad_kernel")) || (rc = cross_copies(C, sb, rb, -1.0f, image, img2))) return rc;
  }
  return KLRTM_OK;
}

int correlator_set_double(Correlator &C, int dz, int dx, int dy, float high_z, float high_xy) {
  KL_REQUIRE((dz == 1 || dz == 2) && (dx == 1 || dx == 2) && (dy == 1 || dy == 2), KLRTM_EINVAL,
             "imaging grid doubling factors must be 1 or 2");
  KL_REQUIRE(dz >= dx && dz >= dy, KLRTM_EINVAL, "a doubled x or y imaging grid needs a doubled z grid too");
  KL_REQUIRE(dy == 1 || C.ny > 1, KLRTM_EINVAL, "no y doubling on a 2D grid");
  // the reference's sub-image layouts: all axes doubled in 3D (interleave3d), z or z and x in 2D (interleave2d
  // walks one x-z plane, bgprtm/ImagingCorrelator.cpp:447-474)
  KL_REQUIRE(C.ny == 1 || (dz == dx && dz == dy), KLRTM_EINVAL,
             "3D imaging grids double all three axes or none (ic_double_z/x/y %d/%d/%d)", dz, dx, dy);
  KL_REQUIRE(high_z > 0.f && high_z <= 1.f && high_xy > 0.f && high_xy <= 1.f, KLRTM_EINVAL, "high-cut fractions in (0, 1]");
  // the new layout is built in locals and committed only once its buffers exist, so a failed allocation
  // leaves the correlator as it was
  const int dbl[3] = {dz, dx, dy};
  int nloop = 1, bits[8] = {0}, src_of[8] = {0}, axis_of[8] = {0}, L[3];
  // copies in the reference's order: z, x, y, zx, zy, xy, zxy (those of doubled axes)
  const int want[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  int idx_of[8];  // copy index of each shift pattern (bits z | x<<1 | y<<2)
  for(int b = 0; b < 8; b++) idx_of[b] = -1;
  idx_of[0] = 0;
  for(int w = 0; w < 7; w++) {
    const int bz = want[w][0], bx = want[w][1], by = want[w][2];
    if((bz && dz < 2) || (bx && dx < 2) || (by && dy < 2)) continue;
    const int b = bz | (bx << 1) | (by << 2);
    // derived from the copy without its last shifted axis (y, then x, then z), shifted along that axis
    const int ax = by ? 2 : bx ? 1 : 0;
    const int from = b & ~(1 << ax);
    bits[nloop] = b, src_of[nloop] = idx_of[from], axis_of[nloop] = ax;
    idx_of[b] = nloop++;
  }
  for(int a = 0; a < 3; a++) {
    const int na = a == 0 ? C.nz : a == 1 ? C.nx : C.ny;
    L[a] = dbl[a] > 1 ? fft_size(na + 16) : 0;
  }
  float *sb = nullptr, *rb = nullptr, *ln = nullptr;
  if(nloop > 1) {
    size_t lmax = 0;
    for(int a = 0; a < 3; a++)
      if(L[a]) lmax = std::max(lmax, C.nxyz / (a == 0 ? C.nz : a == 1 ? C.nx : C.ny) * (size_t)(L[a] + 2));
    hipError_t e = hipMalloc(&sb, sizeof(float) * C.nxyz * nloop);
    if(e == hipSuccess) e = hipMalloc(&rb, sizeof(float) * C.nxyz * nloop);
    if(e == hipSuccess) e = hipMalloc(&ln, sizeof(float) * lmax);
    if(e != hipSuccess) {
      (void)hipFree(sb), (void)hipFree(rb), (void)hipFree(ln);
      set_error("doubled imaging grid: copy buffers of %d volumes: %s", nloop, hipGetErrorString(e));
      return KLRTM_EDEVICE;
    }
  }
  (void)hipFree(C.sbuf), (void)hipFree(C.rbuf), (void)hipFree(C.lines);
  C.sbuf = sb, C.rbuf = rb, C.lines = ln;
  for(int a = 0; a < 3; a++) C.dbl[a] = dbl[a], C.L[a] = L[a];
  C.nloop = nloop;
  for(int i = 0; i < 8; i++) C.bits[i] = bits[i], C.src_of[i] = src_of[i], C.axis_of[i] = axis_of[i];
  C.high_z = high_z, C.high_xy = high_xy;
  return KLRTM_OK;
}

// ImagingCorrelator::interleave: the nloop sub-images of sub onto the doubled grid in out
int correlator_interleave(Correlator &C, const float *sub, float *out) {
  KL_REQUIRE(sub && out && sub != out, KLRTM_EINVAL, "interleave needs distinct input and output");
  if(C.nloop == 1) {
    KL_HIP(hipMemcpyAsync(out, sub, C.nxyz * sizeof(float), hipMemcpyDeviceToDevice, stream()));
    return KLRTM_OK;
  }
  const int4 b03 = make_int4(C.bits[0], C.bits[1], C.bits[2], C.bits[3]);
  const int4 b47 = make_int4(C.bits[4], C.bits[5], C.bits[6], C.bits[7]);
  hipLaunchKernelGGL(corr_interleave_kernel, dim3(nblocks(C.nxyz * C.nloop)), dim3(256), 0, stream(), sub, out, C.nx, C.ny,
                     C.nz, C.dbl[0], C.dbl[1], C.dbl[2], C.nloop, b03, b47);
  return check_launch("corr_interleave_kernel");
}

int correlator_nloop(const Correlator &C) { return C.nloop; }

int correlator_add_illum(Correlator &C, const float *s, float *ill) {
  KL_REQUIRE(s && ill, KLRTM_EINVAL, "null illumination operand");
  hipLaunchKernelGGL(corr_illum_kernel, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), s, ill, C.nxyz);
  return check_launch("corr_illum_kernel");
}

int correlator_interp_illum(Correlator &C, const float *in, float *out) {
  KL_REQUIRE(in && out && in != out, KLRTM_EINVAL, "illumination interpolation needs distinct input and output");
  hipLaunchKernelGGL(corr_interp_illum_kernel, dim3(nblocks(C.nxyz)), dim3(256), 0, stream(), in, out, C.nx, C.ny, C.nz,
                     C.dbl[0], C.dbl[1], C.dbl[2]);
  return check_launch("corr_interp_illum_kernel");
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
  if(C.nloop > 1) {
    // doubled grid: images hold nloop sub-images; applyFFTV's high cut on S and R first
    // (bgprtm/ImagingCondition.cpp:125-134)
    KL_REQUIRE(!image64, KLRTM_EINVAL, "the doubled imaging grid accumulates FP32 images only");
    KL_REQUIRE(option == KLRTM_IC_IMAGE1 || option == KLRTM_IC_INVERSION, KLRTM_EINVAL,
               "the doubled imaging grid takes the IMAGE1 and INVERSION conditions");
    if((rc = hightape(C, src)) || (rc = hightape(C, rec))) return rc;
    if(option == KLRTM_IC_IMAGE1) return cross_term(C, src, rec, 1.0f, image, nullptr);
    if((rc = forward_pair(C, src, rec))) return rc;  // derive (bgprtm/ImagingCorrelator.cpp:376)
    if((rc = cross_term(C, src, rec, 1.0f, image, img_fwi))) return rc;
    return derivative_terms_dbl(C, C.sbuf, C.rbuf, image, img_fwi);
  }
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

// ImagingCorrelator::run on the reference's buffers (after ImagingCondition's applyFFTV, and for IMAGE1 its
// interp): src, rec hold nloop volumes; IMAGE1 correlates the given copies (imaging1,
// bgprtm/ImagingCorrelator.cpp:363-372), INVERSION takes copy 0 and overwrites the copies (inversion1, :374-414)
int correlator_run_copies(Correlator &C, float *image, float *img_fwi, float *src, float *rec, int option) {
  KL_REQUIRE(src && rec, KLRTM_EINVAL, "null wavefield");
  KL_REQUIRE(option == KLRTM_IC_IMAGE1 || option == KLRTM_IC_INVERSION, KLRTM_EINVAL,
             "copies are correlated under IMAGE1 and INVERSION");
  LaunchTimer tm("ic_correlator");
  int rc = KLRTM_OK;
  if(option == KLRTM_IC_IMAGE1) {
    for(int i = 0; i < C.nloop && !rc; i++) {
      const size_t o = (size_t)i * C.nxyz;
      rc = cross_plain(C, src + o, rec + o, 1.0f, image ? image + o : nullptr, nullptr, nullptr, nullptr);
    }
    return rc;
  }
  if((rc = forward_pair(C, src, rec)) || (rc = cross_copies(C, src, rec, 1.0f, image, img_fwi))) return rc;
  return derivative_terms_dbl(C, src, rec, image, img_fwi);
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
int klrtm_correlator_set_double(klrtm_correlator *c, int dz, int dx, int dy, float high_z, float high_xy) {
  KL_REQUIRE(c, KLRTM_EINVAL, "null correlator");
  return correlator_set_double(*reinterpret_cast<Correlator *>(c), dz, dx, dy, high_z, high_xy);
}
int klrtm_correlator_nloop(klrtm_correlator *c, int *nloop) {
  KL_REQUIRE(c && nloop, KLRTM_EINVAL, "null argument");
  *nloop = correlator_nloop(*reinterpret_cast<Correlator *>(c));
  return KLRTM_OK;
}
int klrtm_correlator_interleave(klrtm_correlator *c, const float *sub, float *out) {
  KL_REQUIRE(c, KLRTM_EINVAL, "null correlator");
  return correlator_interleave(*reinterpret_cast<Correlator *>(c), sub, out);
}
int klrtm_correlator_run_copies(klrtm_correlator *c, float *image, float *img_fwi, float *src, float *rec, int option) {
  KL_REQUIRE(c, KLRTM_EINVAL, "null correlator");
  return correlator_run_copies(*reinterpret_cast<Correlator *>(c), image, img_fwi, src, rec, option);
}
int klrtm_correlator_interp(klrtm_correlator *c, float *buf) {
  KL_REQUIRE(c && buf, KLRTM_EINVAL, "null argument");
  return interp(*reinterpret_cast<Correlator *>(c), buf);
}
int klrtm_correlator_hightape(klrtm_correlator *c, float *wave) {
  KL_REQUIRE(c && wave, KLRTM_EINVAL, "null argument");
  return hightape(*reinterpret_cast<Correlator *>(c), wave);
}
int klrtm_correlator_interp_illum(klrtm_correlator *c, const float *ill, float *out) {
  KL_REQUIRE(c, KLRTM_EINVAL, "null correlator");
  return correlator_interp_illum(*reinterpret_cast<Correlator *>(c), ill, out);
}
int klrtm_correlator_kfilter(klrtm_correlator *c, float *wave) {
  KL_REQUIRE(c, KLRTM_EINVAL, "null correlator");
  return correlator_kfilter(*reinterpret_cast<Correlator *>(c), wave);
}
}
