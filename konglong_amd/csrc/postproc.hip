// postproc.hip -- image post-processing after the imaging condition: the
// source-illumination compensation of ImagingCondition::computeDamping /
// applyDamping (bgprtm/ImagingCondition.cpp:600-672), with the box smoothing
// of avgVolume3D / avgVolume1D (bgprtm/volumefilter.cpp:59-135), the cone
// mute of ImagingCondition::imageConeMute (bgprtm/ImagingCondition.cpp:690-756)
// and the Lagrange resampling of the image to the output grid
// (bgprtm/Lagrange.cpp:19-266).
//
// The box filter is a running sum along one axis with edge replication: one
// lane per line keeps the sum and walks the line in the reference's order
// (y lines and x lines: lanes over z, so every step is a coalesced row; z
// lines: one lane per column), so the smoothed volume equals the reference's
// bit for bit.  The mean of the smoothed illumination is a fixed-order FP64
// tree reduction; the damping weights then follow the reference's mixed
// FP32/FP64 arithmetic.  FP contraction is off.
#pragma clang fp contract(off)
#include "klrtm_internal.h"
#include <algorithm>
#include <cmath>
#include <vector>

namespace klrtm {

// box mean of half width h along an axis of n samples with stride s, for every line of the volume
// (lines enumerated by the lane index); out is another buffer
template<int AXIS>
__global__ void box_kernel(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nz, int h,
                           float scaler, float sizea1) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  size_t base, s;
  int n;
  if(AXIS == 2) {  // y lines: lanes over (ix, iz)
    if(c >= (long)nx * nz) return;
    base = (size_t)c, s = (size_t)nx * nz, n = ny;
  } else if(AXIS == 1) {  // x lines: lanes over (iy, iz)
    if(c >= (long)ny * nz) return;
    const int iy = (int)(c / nz), iz = (int)(c % nz);
    base = (size_t)iy * nx * nz + iz, s = (size_t)nz, n = nx;
  } else {  // z lines: one lane per column
    if(c >= (long)nx * ny) return;
    base = (size_t)c * nz, s = 1, n = nz;
  }
  // avgVolume1D: bufm = (h+1)*a[0] + a[1..h]; b[0] = bufm*scaler; then bufm += a[k2] - a[k1]
  float bufm = sizea1 * in[base];
  for(int j = 1; j <= h; j++) bufm += in[base + (size_t)j * s];
  out[base] = bufm * scaler;
  for(int k = 1; k < n; k++) {
    const int k1 = max(0, k - h - 1), k2 = min(n - 1, k + h);
    bufm += (in[base + (size_t)k2 * s] - in[base + (size_t)k1 * s]);
    out[base + (size_t)k * s] = bufm * scaler;
  }
}

__global__ void sum_kernel(const float *__restrict__ v, size_t n, double *__restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    acc += (double)v[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1) {
    if((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if(threadIdx.x == 0) part[blockIdx.x] = red[0];
}

struct DampArgs {
  int nx, ny, nz, ntaper;
  double average, oaverage;
  float damp_weight;
  float wtaper[10];
};

// illum <- 1 / (average * (aa + w*exp(-aa))), aa = illum/average, then the x/y edge taper
__global__ void damping_kernel(float *__restrict__ illum, const DampArgs a) {
  const size_t n = (size_t)a.nx * a.ny * a.nz;
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t col = i / a.nz;
    const int ix = (int)(col % a.nx), iy = (int)(col / a.nx);
    const float aa = (float)((double)illum[i] * a.oaverage);
    const float bb = expf(-aa);
    float v = (float)(1.0 / (a.average * (double)(aa + a.damp_weight * bb)));
    if(iy < a.ntaper && a.ny > 1) v *= a.wtaper[iy];
    if(iy > a.ny - a.ntaper - 1 && a.ny > 1) v *= a.wtaper[a.ny - iy - 1];
    if(ix < a.ntaper) v *= a.wtaper[ix];
    if(ix > a.nx - a.ntaper - 1) v *= a.wtaper[a.nx - ix - 1];
    illum[i] = v;
  }
}

__global__ void multiply_kernel(float *__restrict__ img, const float *__restrict__ w, size_t n) {
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    img[i] = img[i] * w[i];
}

// ImagingCondition::imageConeMute / radius_mute (bgprtm/ImagingCondition.cpp:690-756) on a rectangular
// grid: cells farther than r1(z) laterally from the source are muted with a cosine taper out to 2|r1|;
// r1 = (z - zs)*tan(angle) + offset_min, or linear in a (depth, offset) map
constexpr int MUTE_MAP_MAX = 64;
struct MuteArgs {
  int nx, ny, nz, is_angle, nmap;
  float x0, y0, z0, dx, dy, dz, sx, sy, sz, tangent, offset_min;
  float mz[MUTE_MAP_MAX], mo[MUTE_MAP_MAX];
};
__device__ float radius_mute_dev(const MuteArgs &a, float z) {
  if(a.is_angle) return z * a.tangent + a.offset_min;
  const int n = a.nmap;
  int i = 0;
  if(z > a.mz[0])
    for(i = 0; i < n - 1; i++)
      if(a.mz[i + 1] >= z) break;
  if(i > n - 2) i = n - 2;
  const float z0 = a.mz[i], z1 = a.mz[i + 1];
  return ((z - z0) * a.mo[i + 1] + (z1 - z) * a.mo[i]) / (z1 - z0);
}
__global__ void cone_mute_kernel(float *__restrict__ image, float *__restrict__ weight, const MuteArgs a) {
  const size_t n = (size_t)a.nx * a.ny * a.nz;
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t col = i / a.nz;
    const int iz = (int)(i - col * a.nz), ix = (int)(col % a.nx), iy = (int)(col / a.nx);
    const float mz = (a.z0 + iz * a.dz) - a.sz;
    const float r1 = radius_mute_dev(a, mz);
    const float r2 = 2.0f * fabsf(r1);
    const float my = (a.y0 + iy * a.dy) - a.sy, mx = (a.x0 + ix * a.dx) - a.sx;
    const float rr = sqrtf(mx * mx + my * my);
    float w = 1.0f;
    if(rr > r1) {
      const float d1 = rr - r1, d2 = r2 - r1;
      if(r1 <= 0 || d1 > d2) w = 0.0f;
      else w = 0.5f * (1.0f + cosf((float)(M_PI * (double)d1 / (double)d2)));
    }
    if(image) image[i] *= w;
    if(weight) weight[i] = w;
  }
}

// Lagrange::interpz/x/y (bgprtm/Lagrange.cpp:190-266): out sample = sum_i in[nbeg + i] * coef[i] along one
// axis, i ascending, in FP32 (as applyCoef).  AXIS 0: z (n_in -> n_out along the contiguous axis,
// lines over nx*ny); 1: x (lines over ny, nz); 2: y (lines over nx, nz).
template<int AXIS>
__global__ void lagrange_kernel(const float *__restrict__ in, float *__restrict__ out, int nin, int nout, int na, int nb,
                                const float *__restrict__ coef, const int *__restrict__ norder,
                                const int *__restrict__ nbeg, int np) {
  // output layout: AXIS 0: [na*nb][nout]; 1: [nb = ny][nout][na = nz]; 2: [nout][nb = nx][na = nz]
  const size_t total = (size_t)nout * na * nb;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    int o;
    size_t ibase, s, obase;
    if(AXIS == 0) {
      const size_t line = e / nout;
      o = (int)(e - line * nout);
      ibase = line * nin, s = 1, obase = e;
    } else if(AXIS == 1) {
      const int iz = (int)(e % na);
      const size_t t = e / na;
      o = (int)(t % nout);
      const int iy = (int)(t / nout);
      ibase = (size_t)iy * nin * na + iz, s = (size_t)na, obase = e;
    } else {
      const size_t xz = e % ((size_t)na * nb);
      o = (int)(e / ((size_t)na * nb));
      ibase = xz, s = (size_t)na * nb, obase = e;
    }
    const int nb0 = nbeg[o], no = norder[o];
    float y = 0.0f;
    for(int i = 0; i < no; i++) y += in[ibase + (size_t)(nb0 + i) * s] * coef[(size_t)o * np + i];
    out[obase] = y;
  }
}

static unsigned grid_of(long n) { return (unsigned)std::max(1L, std::min((n + 255) / 256, 256L * 64)); }

// avgVolume3D: y, then x, then z (bgprtm/volumefilter.cpp:59-98)
static int box3d(float *v, float *tmp, int nx, int ny, int nz, int sx, int sy, int sz) {
  auto pass = [&](int axis, const float *in, float *out, int h) {
    const float scaler = (float)(1.0f / ((double)(2.0f * h) + 1.0)), sizea1 = h + 1.0f;
    if(axis == 2)
      hipLaunchKernelGGL(box_kernel<2>, dim3(grid_of((long)nx * nz)), dim3(256), 0, stream(), in, out, nx, ny, nz, h,
                         scaler, sizea1);
    else if(axis == 1)
      hipLaunchKernelGGL(box_kernel<1>, dim3(grid_of((long)ny * nz)), dim3(256), 0, stream(), in, out, nx, ny, nz, h,
                         scaler, sizea1);
    else
      hipLaunchKernelGGL(box_kernel<0>, dim3(grid_of((long)nx * ny)), dim3(256), 0, stream(), in, out, nx, ny, nz, h,
                         scaler, sizea1);
  };
  pass(2, v, tmp, sy);
  pass(1, tmp, v, sx);
  pass(0, v, tmp, sz);
  KL_HIP(hipMemcpyAsync(v, tmp, sizeof(float) * (size_t)nx * ny * nz, hipMemcpyDeviceToDevice, stream()));
  return check_launch("box_kernel");
}

}  // namespace klrtm

using namespace klrtm;

extern "C" {

int klrtm_illum_damping(float *illum, int nx, int ny, int nz, int smooth, int smooth_xy, int smooth_z,
                        float amp_damping, float damp_weight, double *average_out) {
  KL_REQUIRE(illum && nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad illumination volume");
  KL_REQUIRE(!smooth || (smooth_xy >= 0 && smooth_z >= 0 && smooth_xy < nx && smooth_z < nz &&
                         (ny == 1 || smooth_xy < ny)),
             KLRTM_EINVAL, "smoothing half widths %d, %d must be shorter than the axes", smooth_xy, smooth_z);
  const size_t n = (size_t)nx * ny * nz;
  LaunchTimer tm("illum_damping");
  if(smooth) {
    float *tmp = nullptr;
    KL_HIP(hipMallocAsync(&tmp, n * sizeof(float), stream()));
    int rc;
    if(ny > 1) {
      rc = box3d(illum, tmp, nx, ny, nz, smooth_xy, smooth_xy, smooth_z);
    } else {  // avgVolume2D: x then z
      const float sx = (float)(1.0f / ((double)(2.0f * smooth_xy) + 1.0)), ax = smooth_xy + 1.0f;
      const float sz = (float)(1.0f / ((double)(2.0f * smooth_z) + 1.0)), az = smooth_z + 1.0f;
      hipLaunchKernelGGL(box_kernel<1>, dim3(grid_of((long)nz)), dim3(256), 0, stream(), illum, tmp, nx, 1, nz, smooth_xy,
                         sx, ax);
      hipLaunchKernelGGL(box_kernel<0>, dim3(grid_of((long)nx)), dim3(256), 0, stream(), tmp, illum, nx, 1, nz, smooth_z,
                         sz, az);
      rc = check_launch("box_kernel");
    }
    KL_HIP(hipFreeAsync(tmp, stream()));
    if(rc) return rc;
  }
  // mean: per-block FP64 partials in a fixed order, summed on the host in block order
  const unsigned nb = 1024;
  double *part = nullptr;
  KL_HIP(hipMallocAsync(&part, nb * sizeof(double), stream()));
  hipLaunchKernelGGL(sum_kernel, dim3(nb), dim3(256), 0, stream(), illum, n, part);
  std::vector<double> hp(nb);
  KL_HIP(hipMemcpyAsync(hp.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, stream()));
  KL_HIP(hipStreamSynchronize(stream()));
  KL_HIP(hipFreeAsync(part, stream()));
  double total = 0.0;
  for(double v : hp) total += v;
  double average = total / (double)(unsigned long)n;
  average = pow(average, (double)amp_damping);
  KL_REQUIRE(average > 0.0 && std::isfinite(average), KLRTM_EINVAL, "illumination mean %g is not positive", average);
  DampArgs a{};
  a.nx = nx, a.ny = ny, a.nz = nz, a.ntaper = 10;
  a.average = average, a.oaverage = 1.0 / average;
  a.damp_weight = damp_weight;
  for(int i = 0; i < 10; i++) {
    float temp = (float)(M_PI * (double)float(i) / (double)float(10));
    a.wtaper[i] = 0.5f * (1.0f - cosf(temp));
  }
  hipLaunchKernelGGL(damping_kernel, dim3(grid_of((long)n)), dim3(256), 0, stream(), illum, a);
  if(average_out) *average_out = average;
  return check_launch("damping_kernel");
}

int klrtm_avg_volume3d(float *v, int nx, int ny, int nz, int sx, int sy, int sz) {
  KL_REQUIRE(v && nx > 0 && ny > 1 && nz > 0, KLRTM_EINVAL, "bad volume (3D, ny > 1)");
  KL_REQUIRE(sx >= 0 && sy >= 0 && sz >= 0 && sx < nx && sy < ny && sz < nz, KLRTM_EINVAL,
             "half widths %d, %d, %d must be shorter than the axes", sx, sy, sz);
  float *tmp = nullptr;
  KL_HIP(hipMallocAsync(&tmp, (size_t)nx * ny * nz * sizeof(float), stream()));
  const int rc = box3d(v, tmp, nx, ny, nz, sx, sy, sz);
  KL_HIP(hipFreeAsync(tmp, stream()));
  return rc;
}

int klrtm_cone_mute(float *image, float *weight, int nx, int ny, int nz, float x0, float y0, float z0, float dx, float dy,
                    float dz, float sx, float sy, float sz, float angle_deg, float offset_min, const float *map, int nmap) {
  KL_REQUIRE(nx > 0 && ny > 0 && nz > 0, KLRTM_EINVAL, "bad image grid");
  KL_REQUIRE(!map || (nmap >= 2 && nmap <= MUTE_MAP_MAX), KLRTM_EINVAL, "the depth-offset map needs 2..%d pairs",
             MUTE_MAP_MAX);
  MuteArgs a{};
  a.nx = nx, a.ny = ny, a.nz = nz, a.is_angle = map ? 0 : 1, a.nmap = map ? nmap : 0;
  a.x0 = x0, a.y0 = y0, a.z0 = z0, a.dx = dx, a.dy = dy, a.dz = dz, a.sx = sx, a.sy = sy, a.sz = sz;
  a.tangent = tanf((float)(angle_deg * M_PI / 180.0f));
  a.offset_min = offset_min;
  for(int i = 0; map && i < nmap; i++) a.mz[i] = map[2 * i], a.mo[i] = map[2 * i + 1];
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL(cone_mute_kernel, dim3(grid_of((long)n)), dim3(256), 0, stream(), image, weight, a);
  return check_launch("cone_mute_kernel");
}

int klrtm_lagrange_resample(const float *in, int nzi, int nxi, int nyi, float z0i, float dzi, float x0i, float dxi,
                            float y0i, float dyi, float *out, int nzo, int nxo, int nyo, float z0o, float dzo, float x0o,
                            float dxo, float y0o, float dyo, int npoints) {
  KL_REQUIRE(in && out && nzi > 1 && nxi > 0 && nyi > 0 && nzo > 0 && nxo > 0 && nyo > 0, KLRTM_EINVAL, "bad grids");
  KL_REQUIRE(npoints >= 2 && npoints % 2 == 0 && npoints <= 16, KLRTM_EINVAL, "npoints must be even, 2..16");
  const bool do_y = nyo > 1 || nyi > 1, do_x = do_y || nxo > 1;
  KL_REQUIRE(!do_x || nxi > 1, KLRTM_EINVAL, "resampling along x needs nx_in > 1");
  KL_REQUIRE(!do_y || nyi > 1, KLRTM_EINVAL, "resampling along y needs ny_in > 1");
  KL_REQUIRE(do_x || nxo == nxi, KLRTM_EINVAL, "without x resampling nx_out must equal nx_in");
  KL_REQUIRE(do_y || nyo == nyi, KLRTM_EINVAL, "without y resampling ny_out must equal ny_in");
  struct Axis {
    std::vector<float> coef;
    std::vector<int> norder, nbeg;
  };
  // Lagrange::initCoef / checkGrid / getCoef (bgprtm/Lagrange.cpp:121-175); an output sample before the
  // input range takes the first input sample (the reference reads index -1 there), and the stencil is
  // kept inside grids shorter than npoints
  auto init = [&](int nin, float g0i, float gdi, int nout, float g0o, float gdo, Axis &A) {
    std::vector<float> tin(nin), tout(nout);
    for(int i = 0; i < nin; i++) tin[i] = g0i + i * gdi;
    for(int i = 0; i < nout; i++) tout[i] = g0o + i * gdo;
    A.coef.assign((size_t)nout * npoints, 0.f), A.norder.resize(nout), A.nbeg.resize(nout);
    const int mintp = npoints / 2;
    for(int o = 0; o < nout; o++) {
      const float z = tout[o];
      int iz0;
      if(z < tin[0]) iz0 = -1;
      else if(z > tin[nin - 1]) iz0 = nin;
      else {
        iz0 = 0;
        while(z > tin[iz0 + 1] && iz0 <= nin - 1) iz0++;
      }
      int no, nb;
      if(iz0 == -1) no = 1, nb = 0;
      else if(iz0 == nin) no = 1, nb = nin - 1;
      else if(iz0 < mintp) no = (iz0 + 1) * 2, nb = 0;
      else if(nin - 1 - iz0 < mintp) no = (nin - 1 - iz0) * 2, nb = nin - no;
      else no = npoints, nb = iz0 - mintp;
      nb = std::min(std::max(nb, 0), nin - 1), no = std::max(1, std::min(no, nin - nb));  // stay inside tiny grids
      float *c = &A.coef[(size_t)o * npoints];
      for(int i = 0; i < no; i++) {
        c[i] = 1.0f;
        for(int j = 0; j < no; j++)
          if(j != i) c[i] = c[i] * (z - tin[j + nb]) / (tin[i + nb] - tin[j + nb]);
      }
      A.norder[o] = no, A.nbeg[o] = nb;
    }
  };
  Axis az, ax, ay;
  init(nzi, z0i, dzi, nzo, z0o, dzo, az);
  if(do_x) init(nxi, x0i, dxi, nxo, x0o, dxo, ax);
  if(do_y) init(nyi, y0i, dyi, nyo, y0o, dyo, ay);
  std::vector<void *> allocs;
  auto upload = [&](const void *h, size_t bytes) -> void * {
    void *d = nullptr;
    if(hipMallocAsync(&d, std::max<size_t>(bytes, 4), stream()) != hipSuccess) return nullptr;
    allocs.push_back(d);
    if(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream()) != hipSuccess) return nullptr;
    return d;
  };
  int rc = KLRTM_OK;
  auto fail = [&](const char *what) {
    set_error("resampling: %s failed", what);
    rc = KLRTM_ENOMEM;
  };
  const size_t nwz = (size_t)nzo * nxi * nyi, nwx = (size_t)nzo * nxo * nyi;
  float *bz = nullptr, *bx = nullptr;
  if(do_x && hipMallocAsync(&bz, nwz * sizeof(float), stream()) == hipSuccess) allocs.push_back(bz);
  if(do_y && hipMallocAsync(&bx, nwx * sizeof(float), stream()) == hipSuccess) allocs.push_back(bx);
  if((do_x && !bz) || (do_y && !bx)) fail("work allocation");
  const float *cz = rc ? nullptr : (const float *)upload(az.coef.data(), az.coef.size() * 4);
  const int *oz = rc ? nullptr : (const int *)upload(az.norder.data(), az.norder.size() * 4);
  const int *gz = rc ? nullptr : (const int *)upload(az.nbeg.data(), az.nbeg.size() * 4);
  if(!rc && (!cz || !oz || !gz)) fail("coefficient upload");
  LaunchTimer tm("lagrange");
  if(!rc) {
    float *zout = do_x ? bz : out;
    hipLaunchKernelGGL(lagrange_kernel<0>, dim3(grid_of((long)nzo * nxi * nyi)), dim3(256), 0, stream(), in, zout, nzi,
                       nzo, nxi, nyi, cz, oz, gz, npoints);
  }
  if(!rc && do_x) {
    const float *cx = (const float *)upload(ax.coef.data(), ax.coef.size() * 4);
    const int *ox = (const int *)upload(ax.norder.data(), ax.norder.size() * 4);
    const int *gx = (const int *)upload(ax.nbeg.data(), ax.nbeg.size() * 4);
    if(!cx || !ox || !gx) fail("coefficient upload");
    else
      hipLaunchKernelGGL(lagrange_kernel<1>, dim3(grid_of((long)nxo * nzo * nyi)), dim3(256), 0, stream(), bz,
                         do_y ? bx : out, nxi, nxo, nzo, nyi, cx, ox, gx, npoints);
  }
  if(!rc && do_y) {
    const float *cy = (const float *)upload(ay.coef.data(), ay.coef.size() * 4);
    const int *oy = (const int *)upload(ay.norder.data(), ay.norder.size() * 4);
    const int *gy = (const int *)upload(ay.nbeg.data(), ay.nbeg.size() * 4);
    if(!cy || !oy || !gy) fail("coefficient upload");
    else
      hipLaunchKernelGGL(lagrange_kernel<2>, dim3(grid_of((long)nyo * nxo * nzo)), dim3(256), 0, stream(), bx, out, nyi,
                         nyo, nzo, nxo, cy, oy, gy, npoints);
  }
  if(!rc) rc = check_launch("lagrange_kernel");
  KL_HIP(hipStreamSynchronize(stream()));  // the host coefficient tables must outlive their copies
  for(void *d : allocs) (void)hipFreeAsync(d, stream());
  return rc;
}

int klrtm_apply_damping(const float *illum, float *img, size_t n) {
  KL_REQUIRE(illum && img, KLRTM_EINVAL, "null volume");
  hipLaunchKernelGGL(multiply_kernel, dim3(grid_of((long)n)), dim3(256), 0, stream(), img, illum, n);
  return check_launch("multiply_kernel");
}

}  // extern "C"
